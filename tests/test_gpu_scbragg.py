"""Mosaic single-crystal Bragg (SCBragg) on the GPU vs the CPU checker, through
the C ABI.  Tables built at init are bit-identical; oriented cross sections
agree to 1e-12 relative; sampled directions are compared per neutron on the
same Philox stream (a bounded fraction may differ through last-bit libm
differences in acos/cos/sin of the rejection loop)."""
import math

import numpy as np
import pytest

from ncrystal_amd import synth
from conftest import assert_streams_match

pytestmark = pytest.mark.gpu

WL2EKIN = 0.081804209605330899
XS_RTOL = 1e-12
GE_A = 5.65735
GE_D220 = GE_A / math.sqrt(8.0)
GE_220_DEMI = np.array([[2, 2, 0], [2, -2, 0], [2, 0, 2], [2, 0, -2], [0, 2, 2], [0, 2, -2]], dtype=np.int32)


@pytest.fixture(scope="module")
def core():
    from ncrystal_amd import _capi, core
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    return core


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(a), np.abs(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(den > 0, np.abs(a - b) / den, 0.0)


def _pair(core, oracle, a, b, c, v0n, fams, R, mos, prec=1e-3, ntrunc=0.0):
    hkl, mult, dsp, fsq, moff, mh = fams
    L = core.reci_lattice_orthogonal(a, b, c)
    g = core.create_scbragg(v0n, dsp, fsq, moff, mh, L, R, mos, prec, ntrunc)
    o = oracle.scbragg(a, b, c, v0n, dsp, fsq, moff, mh, R, mos, prec, ntrunc)
    return g, o


def _fcc(core, oracle, dcutoff=0.3, rot_seed=4, mos_arcmin=30.0):
    cr = synth.fcc_powder(dcutoff=dcutoff)
    msd = oracle.debye_msd(410.0, 293.0, 26.98)
    sp = [(cr.species[0].csl, msd, cr.species[0].positions)]
    fams = oracle.hkl_orthogonal_members(cr.a, cr.b, cr.c, sp, cr.dcutoff)
    R = synth.random_rotation(rot_seed)
    return _pair(core, oracle, cr.a, cr.b, cr.c, cr.volume * cr.n_atoms, fams, R,
                 mos_arcmin * math.pi / (180 * 60))


def _ge220(core, oracle, mos=math.pi / 180, R=None):
    fams = (GE_220_DEMI[:1], np.array([12]), np.array([GE_D220]), np.array([40.0]), np.array([0, 6]), GE_220_DEMI)
    return _pair(core, oracle, GE_A, GE_A, GE_A, GE_A ** 3 * 8, fams, np.eye(3) if R is None else R, mos)


def _ortho_c5(core, oracle, seed=1, dcutoff=0.2):
    cr = synth.ortho_random(seed=seed, dcutoff=dcutoff)
    sp = [(s.csl, oracle.debye_msd(s.debye_temp, cr.temperature, s.mass), s.positions) for s in cr.species]
    fams = core.hkl_orthogonal_members(cr.a, cr.b, cr.c, sp, cr.dcutoff)
    R = synth.random_rotation(4)
    return _pair(core, oracle, cr.a, cr.b, cr.c, cr.volume * cr.n_atoms, fams, R, 30 * math.pi / (180 * 60))


def test_hkl_members_bit_exact(core, oracle):
    for cr in (synth.fcc_powder(dcutoff=0.3), synth.ortho_random(seed=5, dcutoff=0.6)):
        sp = [(s.csl, oracle.debye_msd(s.debye_temp, cr.temperature, s.mass), s.positions) for s in cr.species]
        g = core.hkl_orthogonal_members(cr.a, cr.b, cr.c, sp, cr.dcutoff)
        o = oracle.hkl_orthogonal_members(cr.a, cr.b, cr.c, sp, cr.dcutoff)
        for i in (0, 1, 2, 4, 5):  # hkl, mult, d, member offsets, members
            assert np.array_equal(g[i], o[i])
        assert rel_err(g[3], o[3]).max() <= 1e-13
        assert np.array_equal(np.diff(g[4]) * 2, g[1])


@pytest.mark.parametrize("case", ["fcc", "ge220", "ortho_c5"])
def test_scbragg_tables_bit_exact(core, oracle, case):
    g, o = {"fcc": _fcc, "ge220": _ge220, "ortho_c5": _ortho_c5}[case](core, oracle)
    gi, oi = g.info(), o.info()
    assert gi["nfam"] == oi["nfam"] and gi["nnormals"] == oi["nnormals"]
    assert gi["threshold"] == oi["threshold"]
    assert np.array_equal(gi["gos"], oi["gos"])
    gt, ot = g.tables(), o.tables()
    for k in ("inv2d", "xsfact", "off", "normals", "lt_evalcosx", "lt_sofcosd"):
        assert np.array_equal(gt[k], ot[k]), k
    print(case, "families", gi["nfam"], "demi-normals", gi["nnormals"])


def test_scbragg_xs_fcc(core, oracle):
    g, o = _fcc(core, oracle)
    n = 100000
    dirs = synth.isotropic_dirs(n, seed=3)
    for wl in (1.0, 2.5, 4.0, 4.6):
        e = np.full(n, WL2EKIN / wl ** 2)
        xg = g.crossSection(e, dirs)
        xo = o.xs_dir(e, dirs, nthreads=8)
        assert rel_err(xg, xo).max() <= XS_RTOL, wl
        assert (xg > 0).any()
    # mixed energies, including below threshold and E = 0
    e = synth.energies_logu(n, 1e-4, 1.0, seed=2)
    e[:10] = 0.0
    xg = g.crossSection(e, dirs)
    xo = o.xs_dir(e, dirs, nthreads=8)
    assert rel_err(xg, xo).max() <= XS_RTOL
    assert np.all(xg[:10] == 0.0)


def _slow_fraction(g, dirs, wl):
    """fraction of neutrons with a contribution outside the fast spline
    range (numpy restatement of the fast-path condition for the one family)"""
    gos = g.info()["gos"]
    cta, k1, k2 = gos[3], gos[7], gos[8]
    nrm = g.tables()["normals"]
    w = math.floor(max(wl, 1e-15) * 1e15 + 0.5) * 1e-15
    i2 = math.floor(max(0.5 / GE_D220, 1e-15) * 1e15 + 0.5) * 1e-15
    ca = w * i2
    sa = math.sqrt(1 - ca * ca)
    slow = np.zeros(len(dirs), dtype=bool)
    for nv in nrm:
        for sgn in (1.0, -1.0):
            cg = np.clip(sgn * dirs @ nv, -1, 1)
            sg = np.sqrt(1 - cg * cg)
            cd = ca * cg + sa * sg
            slow |= (cd > cta) & ~((sa * sg >= 1e-14) & (k2 > k1 * sa * sg + ca * cg))
    return slow.mean()


def test_scbragg_backscattering_slow_path(core, oracle):
    # wavelengths at 0.9995-0.99999 of the 2d threshold and directions near a
    # normal: circle integrals outside the spline validity (Romberg kernels)
    g, o = _ge220(core, oracle)
    nrm = np.array([1.0, 1.0, 0.0]) / math.sqrt(2)
    n = 20000
    dirs = synth.cone_dirs(n, -nrm, 3 * math.pi / 180, seed=7)
    for f in (0.9995, 0.99999):
        wl = 2 * GE_D220 * f
        assert _slow_fraction(g, dirs, wl) > 0.01
        e = np.full(n, WL2EKIN / wl ** 2)
        xg = g.crossSection(e, dirs)
        xo = o.xs_dir(e, dirs, nthreads=8)
        assert rel_err(xg, xo).max() <= XS_RTOL, f
        g.setRandomSeed(5)
        ge, gd = g.sampleScatter(e, dirs)
        oe, od, nerr = o.sample_dir(e, dirs, seed=5, nthreads=8)
        assert nerr == 0 and np.array_equal(ge, e)
        assert_streams_match(np.max(np.abs(gd - od), axis=1) > 1e-9, f"scbragg backscattering {f}")


def test_scbragg_c5_xs_and_sampling(core, oracle):
    g, o = _ortho_c5(core, oracle)
    n = 20000
    wl = synth.wavelengths_uniform(n, 1.0, 6.0, seed=6)
    e = WL2EKIN / wl ** 2
    dirs = synth.cone_dirs(n, synth.random_unit_vector(5), 2 * math.pi / 180, seed=3)
    xg = g.crossSection(e, dirs)
    xo = o.xs_dir(e, dirs, nthreads=8)
    assert rel_err(xg, xo).max() <= XS_RTOL
    assert np.mean(xg > 0) > 0.5
    g.setRandomSeed(11)
    ge, gd = g.sampleScatter(e, dirs)
    oe, od, nerr = o.sample_dir(e, dirs, seed=11, nthreads=8)
    assert nerr == 0 and np.array_equal(ge, e)
    assert np.abs(np.linalg.norm(gd, axis=1) - 1).max() < 1e-14
    assert_streams_match(np.max(np.abs(gd - od), axis=1) > 1e-9, "scbragg c5")
    # Bragg-elastic: the scattering angle of every reflected neutron is 2 theta_B
    # of a family with 2d > wl
    hit = xg > 0
    mu = np.sum(gd * dirs, axis=1)[hit]
    dimp = wl[hit] / np.sqrt(2.0 * (1.0 - mu))  # d implied by the scattering angle
    dd = np.unique(0.5 / g.tables()["inv2d"])
    j = np.clip(np.searchsorted(dd, dimp), 1, len(dd) - 1)
    near = np.minimum(np.abs(dd[j] - dimp), np.abs(dd[j - 1] - dimp)) / dimp
    assert near.max() < 1e-9
    assert np.all(dimp > wl[hit] / 2)


def test_scbragg_isotropic_average_is_powder(core, oracle):
    g, _ = _fcc(core, oracle, dcutoff=0.2)
    cr = synth.fcc_powder(dcutoff=0.2)
    msd = oracle.debye_msd(410.0, 293.0, 26.98)
    hkl, mult, dsp, fsq = oracle.hkl_orthogonal(cr.a, cr.b, cr.c, [(cr.species[0].csl, msd, cr.species[0].positions)],
                                                cr.dcutoff)
    pb = oracle.powderbragg(cr.volume * cr.n_atoms, dsp, fsq * mult)
    n = 2000000
    dirs = synth.isotropic_dirs(n, seed=9)
    for wl in (1.0, 3.0):
        e = WL2EKIN / wl ** 2
        x = g.crossSection(np.full(n, e), dirs)
        ref = pb.xs([e])[0]
        assert abs(x.mean() - ref) < 4 * x.std() / math.sqrt(n) + 2e-3 * ref


def test_scbragg_interface(core, oracle):
    from ncrystal_amd import _capi as C
    g, o = _fcc(core, oracle)
    assert C.call("ncamd_isoriented", g.handle) == 1
    with pytest.raises(C.NCAMDError):
        g.crossSectionIsotropic([0.01])
    # a composition holding it is oriented as well
    comp = core.create_composition([g], [1.0])
    assert C.call("ncamd_isoriented", comp.handle) == 1
    with pytest.raises(C.NCBadInput):
        comp.crossSectionIsotropic([0.01])
    # repeated single (E, dir) form and the device-pointer form
    d = (0.3, -0.5, 0.81)
    g.setRandomSeed(3)
    ge, gd = g.sampleScatter(0.02, d, repeat=5000)
    oe, od, _ = o.sample_dir(np.full(5000, 0.02), np.tile(d, (5000, 1)), seed=3)
    assert_streams_match(np.max(np.abs(gd - od), axis=1) > 1e-9, "scbragg dirs")
    assert np.all(ge == 0.02)
    n = 50000
    e = synth.energies_logu(n, 1e-3, 0.2, seed=4)
    dirs = synth.isotropic_dirs(n, seed=4)
    st = C.Stream()
    de, dd = C.DeviceArray.from_numpy(e), C.DeviceArray.from_numpy(dirs.ravel())
    xs, oe2, od2 = C.DeviceArray(n), C.DeviceArray(n), C.DeviceArray(3 * n)
    g.dev_crossSection(de.ptr, dd.ptr, n, xs.ptr, stream=st.ptr)
    g.setRandomSeed(8)
    g.dev_sampleScatter(de.ptr, dd.ptr, n, oe2.ptr, od2.ptr, stream=st.ptr)
    st.sync()
    assert rel_err(xs.numpy(), o.xs_dir(e, dirs, nthreads=8)).max() <= XS_RTOL
    re_, rd, _ = o.sample_dir(e, dirs, seed=8, nthreads=8)
    assert_streams_match(np.max(np.abs(od2.numpy().reshape(n, 3) - rd), axis=1) > 1e-9, "scbragg device pointers")
