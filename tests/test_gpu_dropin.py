"""The NCrystal C API names (include/ncrystal_dropin.h) on the GPU, bound with
the argument types NCrystal's own ctypes hooks declare for them
(ncrystal_python/src/NCrystal/_chooks.py:148-191,499-577): handles handed over
from ncamd processes, reference counts, casts, names, domains, the batched
and single-neutron calls (identical to the ncamd entry points on the same
streams), the repeat layout out[irepeat*n + i], clones, RNG state strings, the
host-RNG variant and the obsolete genscatter forms."""
import ctypes
import math

import numpy as np
import pytest

from ncrystal_amd import synth

pytestmark = pytest.mark.gpu

_dbl = ctypes.c_double
_dblp = ctypes.POINTER(ctypes.c_double)
_ulong = ctypes.c_ulong
_int = ctypes.c_int
_cstr = ctypes.c_char_p
_voidp = ctypes.c_void_p


class ncrystal_process_t(ctypes.Structure):
    _fields_ = [("internal", ctypes.c_void_p)]


class ncrystal_scatter_t(ctypes.Structure):
    _fields_ = [("internal", ctypes.c_void_p)]


class ncrystal_absorption_t(ctypes.Structure):
    _fields_ = [("internal", ctypes.c_void_p)]


SIGS = {
    # as _chooks.py declares them
    "ncrystal_version_str": (_cstr, ()),
    "ncrystal_sethaltonerror": (_int, (_int,)),
    "ncrystal_setquietonerror": (_int, (_int,)),
    "ncrystal_error": (_int, ()),
    "ncrystal_lasterror": (_cstr, ()),
    "ncrystal_lasterrortype": (_cstr, ()),
    "ncrystal_clearerror": (None, ()),
    "ncrystal_refcount": (_int, (_voidp,)),
    "ncrystal_valid": (_int, (_voidp,)),
    "ncrystal_unref": (None, (_voidp,)),
    "ncrystal_cast_scat2proc": (ncrystal_process_t, (ncrystal_scatter_t,)),
    "ncrystal_cast_abs2proc": (ncrystal_process_t, (ncrystal_absorption_t,)),
    "ncrystal_cast_proc2scat": (ncrystal_scatter_t, (ncrystal_process_t,)),
    "ncrystal_cast_proc2abs": (ncrystal_absorption_t, (ncrystal_process_t,)),
    "ncrystal_name": (_cstr, (ncrystal_process_t,)),
    "ncrystal_isnonoriented": (_int, (ncrystal_process_t,)),
    "ncrystal_crosssection_nonoriented": (None, (ncrystal_process_t, _dbl, _dblp)),
    "ncrystal_crosssection_nonoriented_many": (None, (ncrystal_process_t, _dblp, _ulong, _ulong, _dblp)),
    "ncrystal_domain": (None, (ncrystal_process_t, _dblp, _dblp)),
    "ncrystal_samplescatterisotropic": (None, (ncrystal_scatter_t, _dbl, _dblp, _dblp)),
    "ncrystal_samplescatterisotropic_many": (None, (ncrystal_scatter_t, _dblp, _ulong, _ulong, _dblp, _dblp)),
    "ncrystal_samplescatter": (None, (ncrystal_scatter_t, _dbl, _dbl * 3, _dblp, _dbl * 3)),
    "ncrystal_samplescatter_many": (None, (ncrystal_scatter_t, _dbl, _dbl * 3, _ulong, _dblp, _dblp, _dblp, _dblp)),
    "ncrystal_crosssection": (None, (ncrystal_process_t, _dbl, _dbl * 3, _dblp)),
    "ncrystal_genscatter_nonoriented": (None, (ncrystal_scatter_t, _dbl, _dblp, _dblp)),
    "ncrystal_genscatter_nonoriented_many": (None, (ncrystal_scatter_t, _dblp, _ulong, _ulong, _dblp, _dblp)),
    "ncrystal_clone_scatter": (ncrystal_scatter_t, (ncrystal_scatter_t,)),
    "ncrystal_clone_scatter_rngbyidx": (ncrystal_scatter_t, (ncrystal_scatter_t, _ulong)),
    "ncrystal_getrngstate_ofscatter": (_voidp, (ncrystal_scatter_t,)),
    "ncrystal_setrngstate_ofscatter": (None, (ncrystal_scatter_t, _cstr)),
    "ncrystal_dealloc_string": (None, (_voidp,)),
}


@pytest.fixture(scope="module")
def nc():
    from ncrystal_amd import _capi, core, materials
    L = _capi.lib()
    if L.ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    lib = ctypes.CDLL(str(_capi.LIBPATH))
    for name, (res, args) in SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib.ncamd_to_ncrystal_scatter.restype = ncrystal_scatter_t
    lib.ncamd_to_ncrystal_scatter.argtypes = [_capi.Handle]
    lib.ncamd_to_ncrystal_absorption.restype = ncrystal_absorption_t
    lib.ncamd_to_ncrystal_absorption.argtypes = [_capi.Handle]
    lib.ncrystal_sethaltonerror(0)
    lib.ncrystal_setquietonerror(1)
    return lib, core, materials


def _arr(n):
    a = np.empty(n)
    return a, a.ctypes.data_as(_dblp)


def _state(lib, s):
    p = lib.ncrystal_getrngstate_ofscatter(s)
    txt = ctypes.cast(p, _cstr).value.decode()
    lib.ncrystal_dealloc_string(p)
    return txt


def test_isotropic_material_through_dropin(nc):
    lib, core, materials = nc
    assert lib.ncrystal_version_str() == b"4.4.2"
    cr = synth.fcc_powder()
    m = materials.Material(cr, sab=synth.small_sab_table(mass_amu=26.98))
    g = m.scatter
    s = lib.ncamd_to_ncrystal_scatter(g.handle)
    assert lib.ncrystal_error() == 0 and s.internal
    assert lib.ncrystal_refcount(ctypes.byref(s)) == g.refCount() == 2
    p = lib.ncrystal_cast_scat2proc(s)
    assert lib.ncrystal_name(p) == b"ProcComposition"
    assert lib.ncrystal_isnonoriented(p) == 1
    lo, hi = _dbl(), _dbl()
    lib.ncrystal_domain(p, lo, hi)
    assert (lo.value, hi.value) == (0.0, math.inf)
    assert not lib.ncrystal_cast_proc2abs(p).internal and lib.ncrystal_cast_proc2scat(p).internal == s.internal
    # batched cross sections: layout out[irepeat*n + i], identical to ncamd
    e = synth.energies_logu(5000, 1e-5, 10.0, seed=3)
    xs, xs_p = _arr(3 * e.size)
    lib.ncrystal_crosssection_nonoriented_many(p, e.ctypes.data_as(_dblp), e.size, 3, xs_p)
    assert lib.ncrystal_error() == 0
    ref = g.crossSectionIsotropic(e)
    assert np.array_equal(xs.reshape(3, -1), np.tile(ref, (3, 1)))
    one = _dbl()
    lib.ncrystal_crosssection_nonoriented(p, float(e[7]), one)
    assert one.value == ref[7]
    lib.ncrystal_crosssection(p, float(e[9]), (_dbl * 3)(0.0, 0.6, 0.8), one)
    assert one.value == ref[9]
    # sampling on the handle's stream: same as the ncamd call with the same state
    lib.ncrystal_setrngstate_ofscatter(s, b"77:0")
    assert _state(lib, s) == "77:0" and g.getRNGState() == (77, 0)
    ef, ef_p = _arr(2 * e.size)
    mu, mu_p = _arr(2 * e.size)
    lib.ncrystal_samplescatterisotropic_many(s, e.ctypes.data_as(_dblp), e.size, 2, ef_p, mu_p)
    assert lib.ncrystal_error() == 0 and _state(lib, s) == "77:1"
    g.setRNGState(77, 0)
    ge, gm = g.sampleScatterIsotropic(e, repeat=2)
    assert np.array_equal(ef, ge.ravel()) and np.array_equal(mu, gm.ravel())
    # obsolete angle / energy-change form on the same stream
    lib.ncrystal_setrngstate_ofscatter(s, b"77:0")
    ang, ang_p = _arr(2 * e.size)
    de, de_p = _arr(2 * e.size)
    lib.ncrystal_genscatter_nonoriented_many(s, e.ctypes.data_as(_dblp), e.size, 2, ang_p, de_p)
    # acos of the C library and numpy's may differ in the last bit
    assert np.max(np.abs(ang - np.arccos(mu))) <= 8e-16 and np.array_equal(de, ef - np.tile(e, 2))
    # one energy and direction, repeated: unit directions
    n = 4000
    oe, oe_p = _arr(n)
    ux, ux_p = _arr(n)
    uy, uy_p = _arr(n)
    uz, uz_p = _arr(n)
    lib.ncrystal_samplescatter_many(s, 0.025, (_dbl * 3)(0.0, 0.0, 1.0), n, oe_p, ux_p, uy_p, uz_p)
    assert lib.ncrystal_error() == 0
    assert np.abs(np.sqrt(ux * ux + uy * uy + uz * uz) - 1).max() < 1e-12 and np.all(oe >= 0)
    # single neutrons
    e1, m1 = _dbl(), _dbl()
    lib.ncrystal_samplescatterisotropic(s, 0.025, e1, m1)
    d1 = (_dbl * 3)()
    lib.ncrystal_samplescatter(s, 0.025, (_dbl * 3)(1.0, 0.0, 0.0), e1, d1)
    assert lib.ncrystal_error() == 0 and abs(math.sqrt(sum(x * x for x in d1)) - 1) < 1e-12
    # release: the object lives on through the ncamd handle
    lib.ncrystal_unref(ctypes.byref(s))
    assert g.refCount() == 1 and s.internal


def test_clones_and_host_rng(nc):
    lib, core, materials = nc
    cr = synth.fcc_powder()
    m = materials.Material(cr)
    g = m.scatter
    g.setRandomSeed(5)
    s = lib.ncamd_to_ncrystal_scatter(g.handle)
    c1 = lib.ncrystal_clone_scatter(s)
    c2 = lib.ncrystal_clone_scatter_rngbyidx(s, 3)
    c3 = lib.ncrystal_clone_scatter_rngbyidx(s, 3)
    assert lib.ncrystal_error() == 0 and c1.internal and c2.internal
    assert lib.ncrystal_refcount(ctypes.byref(c1)) == 1
    e = synth.energies_logu(20000, 1e-4, 1.0, seed=2)
    ep = e.ctypes.data_as(_dblp)
    xa, xa_p = _arr(e.size)
    xb, xb_p = _arr(e.size)
    lib.ncrystal_crosssection_nonoriented_many(lib.ncrystal_cast_scat2proc(s), ep, e.size, 1, xa_p)
    lib.ncrystal_crosssection_nonoriented_many(lib.ncrystal_cast_scat2proc(c1), ep, e.size, 1, xb_p)
    assert np.array_equal(xa, xb)  # shared tables

    def draw(h):
        a, ap = _arr(e.size)
        b, bp = _arr(e.size)
        lib.ncrystal_samplescatterisotropic_many(h, ep, e.size, 1, ap, bp)
        assert lib.ncrystal_error() == 0
        return b

    m0, m1, m2, m3 = draw(s), draw(c1), draw(c2), draw(c3)
    assert not np.array_equal(m0, m1) and not np.array_equal(m1, m2)
    assert np.array_equal(m2, m3)  # equal stream index, equal stream
    # a clone keeps the source alive; releasing everything deletes it
    for c in (c1, c2, c3):
        lib.ncrystal_unref(ctypes.byref(c))
        assert not c.internal
    assert g.refCount() == 2
    # host-RNG variant: two callback draws key the neutron's stream; the
    # handle's own stream is untouched
    calls = []
    rng = np.random.default_rng(1)
    CB = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p)

    @CB
    def cb(_state):
        calls.append(1)
        return float(rng.uniform())

    f = lib.ncrystal_samplescatter_rs
    f.restype = None
    f.argtypes = (CB, _voidp, ncrystal_scatter_t, _dbl, _dbl * 3, _dblp, _dbl * 3)
    before = _state(lib, s)
    ef, d = _dbl(), (_dbl * 3)()
    for _ in range(50):
        f(cb, None, s, 0.05, (_dbl * 3)(0.0, 0.0, 1.0), ef, d)
        assert lib.ncrystal_error() == 0 and abs(math.sqrt(sum(x * x for x in d)) - 1) < 1e-12
    assert len(calls) == 100 and _state(lib, s) == before
    lib.ncrystal_unref(ctypes.byref(s))


def test_oriented_and_absorption_handles(nc):
    lib, core, materials = nc
    from test_gpu_scbragg import _fcc
    import pyoracle
    g, _ = _fcc(core, pyoracle)
    s = lib.ncamd_to_ncrystal_scatter(g.handle)
    p = lib.ncrystal_cast_scat2proc(s)
    assert lib.ncrystal_name(p) == b"SCBragg" and lib.ncrystal_isnonoriented(p) == 0
    lo, hi = _dbl(), _dbl()
    lib.ncrystal_domain(p, lo, hi)
    assert lo.value == g.info()["threshold"] and hi.value == math.inf
    # the isotropic calls are refused with poisoned outputs
    x = _dbl()
    lib.ncrystal_crosssection_nonoriented(p, 0.05, x)
    assert x.value == -1.0 and lib.ncrystal_error() == 1
    assert lib.ncrystal_lasterrortype() == b"BadInput"
    lib.ncrystal_clearerror()
    lib.ncrystal_unref(ctypes.byref(s))
    a = core.create_absoov(0.231)
    ha = lib.ncamd_to_ncrystal_absorption(a.handle)
    pa = lib.ncrystal_cast_abs2proc(ha)
    assert lib.ncrystal_name(pa) == b"AbsOOV"
    assert not lib.ncrystal_cast_proc2scat(pa).internal and lib.ncrystal_cast_proc2abs(pa).internal
    lib.ncrystal_crosssection_nonoriented(pa, 0.0253, x)
    assert x.value == pytest.approx(0.231 * math.sqrt(0.0252998 / 0.0253), rel=1e-4)
    lib.ncrystal_unref(ctypes.byref(ha))
    # a scatter handle cannot wrap an absorption process
    bad = lib.ncamd_to_ncrystal_scatter(a.handle)
    assert not bad.internal and lib.ncrystal_error() == 1
    lib.ncrystal_clearerror()
