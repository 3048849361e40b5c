"""ctypes front end of the CPU checker (oracle/ncoracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py``.  The product package never imports it.
"""
import ctypes
import os
import pathlib

import math

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_LIBPATH = _HERE / "build" / "libncoracle.so"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not _LIBPATH.exists():
            build()
        L = ctypes.CDLL(str(_LIBPATH))
        L.orc_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        L.orc_u64_to_fp01.argtypes = [ctypes.c_uint64]
        L.orc_u64_to_fp01.restype = ctypes.c_double
        L.orc_rng_fill.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, _dp]
        L.orc_integrate_alpha_interval_fast.argtypes = [ctypes.c_double] * 6
        L.orc_integrate_alpha_interval_fast.restype = ctypes.c_double
        L.orc_sincos_2pix.argtypes = [ctypes.c_double, _dp, _dp]
        L.orc_debye_msd.argtypes = [ctypes.c_double] * 3
        L.orc_debye_msd.restype = ctypes.c_double
        L.orc_alpha_limits.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_freegas_xs_shape.argtypes = [ctypes.c_double]
        L.orc_freegas_xs_shape.restype = ctypes.c_double
        L.orc_hkl_orthogonal.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int, _ip, _dp, _dp, _dp,
                                                                  ctypes.c_double, ctypes.c_double,
                                                                  ctypes.c_double, ctypes.c_int, _ip, _ip,
                                                                  _dp, _dp]
        L.orc_hkl_orthogonal.restype = ctypes.c_int
        L.orc_hkl_orthogonal_members.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int, _ip, _dp, _dp, _dp,
                                                                          ctypes.c_double, ctypes.c_double,
                                                                          ctypes.c_double, ctypes.c_int, _ip, _ip,
                                                                          _dp, _dp, _i64p, _ip, ctypes.c_int64]
        L.orc_hkl_orthogonal_members.restype = ctypes.c_int
        L.orc_hkl_cell_members.argtypes = [ctypes.c_double] * 6 + [ctypes.c_int, _ip, _dp, _dp, _dp, ctypes.c_double,
                                                                    ctypes.c_double, ctypes.c_double, ctypes.c_int, _ip,
                                                                    _ip, _dp, _dp, _i64p, _ip, ctypes.c_int64]
        L.orc_hkl_cell_members.restype = ctypes.c_int
        L.orc_debug_sort_fams.argtypes = [ctypes.c_int, _dp, _dp, _ip, _ip, _ip]
        L.orc_hkl_spacegroup_members.argtypes = ([ctypes.c_int] + [ctypes.c_double] * 6 +
                                                 [ctypes.c_int, _ip, _dp, _dp, _dp, ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_int, _ip, _ip, _dp, _dp, _i64p, _ip,
                                                  ctypes.c_int64])
        L.orc_hkl_spacegroup_members.restype = ctypes.c_int
        L.orc_reci_lattice.argtypes = [ctypes.c_double] * 6 + [_dp]
        L.orc_create_freegas.argtypes = [ctypes.c_double] * 3
        L.orc_create_freegas.restype = _vp
        L.orc_set_freegas_procedure.argtypes = [ctypes.c_int]
        L.orc_set_freegas_procedure.restype = None
        L.orc_create_absoov.argtypes = [ctypes.c_double]
        L.orc_create_absoov.restype = _vp
        L.orc_reci_lattice.restype = ctypes.c_int
        L.orc_create_scbragg.argtypes = [_dp, ctypes.c_double, ctypes.c_int, _dp, _dp, _i64p, _ip, _dp,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_create_scbragg.restype = _vp
        L.orc_scb_info.argtypes = [_vp, _ip, _i64p, _dp, _dp]
        L.orc_scb_rawxs.argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.orc_scb_rawxs.restype = None
        L.orc_scb_tables.argtypes = [_vp, _dp, _dp, _i64p, _dp, _dp, _dp]
        L.orc_xs_dir_many.argtypes = [_vp, _dp, _dp, ctypes.c_int64, _dp, ctypes.c_int]
        L.orc_create_powderbragg.argtypes = [ctypes.c_double, _dp, _dp, ctypes.c_int]
        L.orc_create_powderbragg.restype = _vp
        L.orc_pb_ntables.argtypes = [_vp]
        L.orc_pb_tables.argtypes = [_vp, _dp, _dp]
        L.orc_create_elinc.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_create_elinc.restype = _vp
        L.orc_create_sab.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_create_sab.restype = _vp
        L.orc_sab_negrid.argtypes = [_vp]
        L.orc_sab_egrid.argtypes = [_vp, _dp, _dp]
        L.orc_sab_sampler_info.argtypes = [_vp, ctypes.c_int, _ip, _ip, _ip, _dp]
        L.orc_sab_sampler_arrays.argtypes = [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip]
        L.orc_sab_derived.argtypes = [_vp, _dp, _dp]
        L.orc_sab_analyse_xs.argtypes = [_vp, ctypes.c_double]
        L.orc_sab_analyse_xs.restype = ctypes.c_double
        L.orc_create_composition.argtypes = [ctypes.c_int, ctypes.POINTER(_vp), _dp]
        L.orc_create_composition.restype = _vp
        L.orc_free.argtypes = [_vp]
        L.orc_comp_ncomponents.argtypes = [_vp]
        L.orc_comp_component.argtypes = [_vp, ctypes.c_int, _dp]
        L.orc_comp_component_ptr.argtypes = [_vp, ctypes.c_int]
        L.orc_comp_component_ptr.restype = _vp
        L.orc_elinc_n.argtypes = [_vp]
        L.orc_elinc_data.argtypes = [_vp, _dp, _dp]
        L.orc_xs_many.argtypes = [_vp, _dp, ctypes.c_int64, _dp, ctypes.c_int]
        L.orc_sample_iso_many.argtypes = [_vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, _dp,
                                          ctypes.c_int64, ctypes.c_int64, _dp, _dp, ctypes.c_int]
        L.orc_sample_iso_many.restype = ctypes.c_int64
        L.orc_sab_sample_ab_many.argtypes = [_vp, ctypes.c_uint64, ctypes.c_uint32, _dp, ctypes.c_int64,
                                             _dp, _dp, ctypes.c_int]
        L.orc_sab_sample_ab_many.restype = ctypes.c_int64
        L.orc_set_alg1_variant.argtypes = [ctypes.c_int]
        L.orc_set_alg1_variant.restype = None
        L.orc_pwdist_percentiles.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp,
                                             ctypes.POINTER(ctypes.c_uint)]
        L.orc_pwdist_percentiles.restype = ctypes.c_int
        L.orc_tailed_breakdown.argtypes = [_dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_uint, ctypes.c_uint, _dp, ctypes.POINTER(ctypes.c_uint)]
        L.orc_tailed_breakdown.restype = None
        L.orc_sab_sample_alpha.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_sab_sample_alpha.restype = None
        L.orc_sab_sample_highe_many.argtypes = [_vp, ctypes.c_uint64, ctypes.c_uint32, _dp, ctypes.c_int64, _dp, _dp]
        L.orc_sab_sample_highe_many.restype = ctypes.c_int64
        L.orc_sab_std_format.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _dp, _dp, _dp,
                                         ctypes.POINTER(ctypes.c_int64)]
        L.orc_sab_std_format.restype = ctypes.c_int
        L.orc_vdos_lasterror.restype = ctypes.c_char_p
        L.orc_vdos_regularise.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
        L.orc_vdos_debye.argtypes = [ctypes.c_double, _dp, _dp, _dp]
        L.orc_vdos_debye.restype = None
        L.orc_vdos_analyse.argtypes = [ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, _dp]
        L.orc_vdos_gn.argtypes = [ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
        L.orc_vdos_kernel_create.argtypes = [ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_uint, ctypes.c_double]
        L.orc_vdos_kernel_create.restype = _vp
        L.orc_vdos_kernel_dims.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_int), _dp]
        L.orc_vdos_kernel_copy.argtypes = [_vp, _dp, _dp, _dp]
        L.orc_vdos_kernel_copy.restype = None
        L.orc_vdos_kernel_free.argtypes = [_vp]
        L.orc_vdos_kernel_free.restype = None
        L.orc_sample_dir_many.argtypes = [_vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, _dp, _dp,
                                          ctypes.c_int64, _dp, _dp, ctypes.c_int]
        L.orc_sample_dir_many.restype = ctypes.c_int64
        _lib = L
    return _lib


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def rng_fill(seed, stream, idx, n):
    out = np.empty(n)
    lib().orc_rng_fill(seed, stream, idx, n, _d(out))
    return out


def debye_msd(debye_temp, temp, mass):
    return lib().orc_debye_msd(debye_temp, temp, mass)


def alpha_limits(ediv, beta):
    a = ctypes.c_double()
    b = ctypes.c_double()
    lib().orc_alpha_limits(ediv, beta, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def hkl_spacegroup_members(spacegroup, cell, species, dcutoff, fsquarecut=1e-5, maxout=2000000, maxmem=8000000):
    """cell = (a, b, c, alpha, beta, gamma), angles in degrees; species as for
    hkl_orthogonal.  Returns (hkl, mult, dsp, fsq, mem_off, mem_hkl)."""
    natoms = np.array([len(s[2]) for s in species], dtype=np.int32)
    pos = np.ascontiguousarray(np.concatenate([np.asarray(s[2], dtype=np.float64).reshape(-1, 3)
                                               for s in species]).ravel())
    csl = np.array([s[0] for s in species], dtype=np.float64)
    msd = np.array([s[1] for s in species], dtype=np.float64)
    hkl = np.empty(3 * maxout, dtype=np.int32)
    mult = np.empty(maxout, dtype=np.int32)
    dsp = np.empty(maxout)
    fsq = np.empty(maxout)
    moff = np.empty(maxout + 1, dtype=np.int64)
    mhkl = np.empty(3 * maxmem, dtype=np.int32)
    n = lib().orc_hkl_spacegroup_members(int(spacegroup), *[float(x) for x in cell], len(species), _i(natoms),
                                         _d(pos), _d(csl), _d(msd), dcutoff, fsquarecut, maxout, _i(hkl), _i(mult),
                                         _d(dsp), _d(fsq), moff.ctypes.data_as(_i64p), _i(mhkl), maxmem)
    if n < 0:
        raise RuntimeError(f"oracle space-group hkl enumeration failed ({n})")
    nm = int(moff[n])
    return (hkl[:3 * n].reshape(n, 3).copy(), mult[:n].copy(), dsp[:n].copy(), fsq[:n].copy(),
            moff[:n + 1].copy(), mhkl[:3 * nm].reshape(nm, 3).copy())


def reci_lattice(cell):
    m = np.empty(9)
    if lib().orc_reci_lattice(*[float(x) for x in cell], _d(m)):
        raise RuntimeError("invalid cell")
    return m


def hkl_orthogonal(a, b, c, species, dcutoff, fsquarecut=1e-5, merge_tolerance=1e-6, maxout=2000000):
    """species: list of (csl_sqrtbarn, msd, positions[n,3])."""
    natoms = np.array([len(s[2]) for s in species], dtype=np.int32)
    pos = np.ascontiguousarray(np.concatenate([np.asarray(s[2], dtype=np.float64).reshape(-1, 3)
                                               for s in species]).ravel())
    csl = np.array([s[0] for s in species], dtype=np.float64)
    msd = np.array([s[1] for s in species], dtype=np.float64)
    hkl = np.empty(3 * maxout, dtype=np.int32)
    mult = np.empty(maxout, dtype=np.int32)
    dsp = np.empty(maxout)
    fsq = np.empty(maxout)
    n = lib().orc_hkl_orthogonal(a, b, c, len(species), _i(natoms), _d(pos), _d(csl), _d(msd), dcutoff,
                                 fsquarecut, merge_tolerance, maxout, _i(hkl), _i(mult), _d(dsp), _d(fsq))
    if n < 0:
        raise RuntimeError(f"oracle hkl enumeration failed ({n})")
    return hkl[:3 * n].reshape(n, 3).copy(), mult[:n].copy(), dsp[:n].copy(), fsq[:n].copy()


class Proc:
    def __init__(self, ptr, keep=()):
        if not ptr:
            raise RuntimeError("oracle process creation failed")
        self.ptr = ptr
        self._keep = keep

    def __del__(self):
        try:
            if self.ptr and not getattr(self, "_borrowed", False):
                lib().orc_free(self.ptr)
        except Exception:
            pass

    def xs(self, ekin, nthreads=1):
        e = np.ascontiguousarray(np.atleast_1d(np.asarray(ekin, dtype=np.float64)))
        out = np.empty_like(e)
        lib().orc_xs_many(self.ptr, _d(e), e.size, _d(out), nthreads)
        return out

    def sample_iso(self, ekin, repeat=1, seed=1, stream=0, idx_offset=0, nthreads=1):
        e = np.ascontiguousarray(np.atleast_1d(np.asarray(ekin, dtype=np.float64)))
        oe = np.empty(e.size * repeat)
        om = np.empty(e.size * repeat)
        nerr = lib().orc_sample_iso_many(self.ptr, seed, stream, idx_offset, _d(e), e.size, repeat,
                                         _d(oe), _d(om), nthreads)
        return oe, om, nerr

    def xs_dir(self, ekin, dirs, nthreads=1):
        """cross sections with neutron directions (SCBragg and compositions
        holding it; the direction does not enter for isotropic processes)"""
        e = np.ascontiguousarray(np.atleast_1d(np.asarray(ekin, dtype=np.float64)))
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dirs, dtype=np.float64), (e.size, 3)))
        out = np.empty(e.size)
        lib().orc_xs_dir_many(self.ptr, _d(e), _d(d), e.size, _d(out), nthreads)
        return out

    def sample_dir(self, ekin, dirs, seed=1, stream=0, idx_offset=0, nthreads=1):
        e = np.ascontiguousarray(np.atleast_1d(np.asarray(ekin, dtype=np.float64)))
        d = np.ascontiguousarray(np.asarray(dirs, dtype=np.float64).reshape(-1, 3))
        oe = np.empty(e.size)
        od = np.empty((e.size, 3))
        nerr = lib().orc_sample_dir_many(self.ptr, seed, stream, idx_offset, _d(e), _d(d), e.size,
                                         _d(oe), _d(od), nthreads)
        return oe, od, nerr


def freegas(temperature, mass_amu, sigma_free):
    return Proc(lib().orc_create_freegas(temperature, mass_amu, sigma_free))


def pwdist_percentiles(x, y, p):
    """PointwiseDist percentiles and bin indices of the piecewise-linear
    density through (x, y)"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.empty(p.size)
    idx = np.empty(p.size, dtype=np.uint32)
    rc = lib().orc_pwdist_percentiles(x.size, _d(x), _d(y), p.size, _d(p), _d(out),
                                      idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)))
    if rc:
        raise ValueError("invalid pointwise distribution")
    return out, idx


def sab_std_format(knltype, alpha, beta, sab, beta_grid_optimised=False):
    """transformKernelToStdFormat: dict(alpha, beta, sab[nb][na], ntrimmed)"""
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    s = np.ascontiguousarray(sab, dtype=np.float64).reshape(b.size, a.size)
    na, nb, nt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
    args = (int(knltype), int(bool(beta_grid_optimised)), a.size, b.size, _d(a), _d(b), _d(s))
    if lib().orc_sab_std_format(*args, ctypes.byref(na), ctypes.byref(nb), None, None, None, ctypes.byref(nt)):
        raise ValueError("invalid scattering kernel")
    oa, ob, os_ = np.empty(na.value), np.empty(nb.value), np.empty((nb.value, na.value))
    lib().orc_sab_std_format(*args, ctypes.byref(na), ctypes.byref(nb), _d(oa), _d(ob), _d(os_), ctypes.byref(nt))
    return dict(alpha=oa, beta=ob, sab=os_, ntrimmed=nt.value)


def _vdos_error():
    msg = lib().orc_vdos_lasterror()
    return RuntimeError(msg.decode() if msg else "VDOS error")


def vdos_regularise(egrid, density):
    """regulariseVDOSGrid: (emin, emax, density) on the regular grid"""
    eg = np.ascontiguousarray(egrid, dtype=np.float64)
    d = np.ascontiguousarray(density, dtype=np.float64)
    lo, hi = ctypes.c_double(), ctypes.c_double()
    n = lib().orc_vdos_regularise(_d(eg), eg.size, _d(d), d.size, ctypes.byref(lo), ctypes.byref(hi), None, 0)
    if n < 0:
        raise _vdos_error()
    out = np.empty(n)
    lib().orc_vdos_regularise(_d(eg), eg.size, _d(d), d.size, ctypes.byref(lo), ctypes.byref(hi), _d(out), n)
    return lo.value, hi.value, out


def vdos_debye(debye_temp):
    """createVDOSDebye: (emin, emax, density of 20 points)"""
    lo, hi, d = ctypes.c_double(), ctypes.c_double(), np.empty(20)
    lib().orc_vdos_debye(float(debye_temp), ctypes.byref(lo), ctypes.byref(hi), _d(d))
    return lo.value, hi.value, d


def vdos_analyse(emin, emax, density, temperature, mass):
    """dict(gamma0, msd [Aa^2], teff [K], integral) of a regularised VDOS"""
    d = np.ascontiguousarray(density, dtype=np.float64)
    out = np.empty(4)
    if lib().orc_vdos_analyse(emin, emax, _d(d), d.size, temperature, mass, _d(out)):
        raise _vdos_error()
    return dict(gamma0=out[0], msd=out[1], teff=out[2], integral=out[3])


def vdos_gn(emin, emax, density, temperature, mass, order):
    """(egrid, spectrum) of the phonon-order-n function G_n"""
    d = np.ascontiguousarray(density, dtype=np.float64)
    lo, bw = ctypes.c_double(), ctypes.c_double()
    args = (emin, emax, _d(d), d.size, temperature, mass, int(order))
    n = lib().orc_vdos_gn(*args, ctypes.byref(lo), ctypes.byref(bw), None, 0)
    if n < 0:
        raise _vdos_error()
    spec = np.empty(n)
    lib().orc_vdos_gn(*args, ctypes.byref(lo), ctypes.byref(bw), _d(spec), n)
    return lo.value + bw.value * np.arange(n), spec


def vdos_kernel(emin, emax, density, temperature, mass, vdoslux=3, target_emax=0.0):
    """createScatteringKernel: dict(alpha, beta, sab[nb][na], suggested_emax,
    msd, gamma0, norders)"""
    d = np.ascontiguousarray(density, dtype=np.float64)
    k = lib().orc_vdos_kernel_create(emin, emax, _d(d), d.size, temperature, mass, int(vdoslux), float(target_emax))
    if not k:
        raise _vdos_error()
    try:
        na, nb, no, info = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), np.empty(3)
        lib().orc_vdos_kernel_dims(k, ctypes.byref(na), ctypes.byref(nb), ctypes.byref(no), _d(info))
        a, b, s_ = np.empty(na.value), np.empty(nb.value), np.empty((nb.value, na.value))
        lib().orc_vdos_kernel_copy(k, _d(a), _d(b), _d(s_))
    finally:
        lib().orc_vdos_kernel_free(k)
    return dict(alpha=a, beta=b, sab=s_, suggested_emax=info[0], msd=info[1], gamma0=info[2], norders=no.value)


def tailed_breakdown(alpha, sab, logsab, cumul, alow, aupp, ilo, iup):
    """createTailedBreakdown on one slice: (xs_front, xs_middle, xs_back)"""
    a, s, ls, c = (np.ascontiguousarray(v, dtype=np.float64) for v in (alpha, sab, logsab, cumul))
    out = np.empty(3)
    imid = (ctypes.c_uint * 2)()
    lib().orc_tailed_breakdown(_d(a), a.size, _d(s), _d(ls), _d(c), alow, aupp, ilo, iup, _d(out), imid)
    return out


class alg1_variant:
    """Context manager planting a change in the checker's Alg1 procedure
    (1 no first-bin rule, 2 egrid margin 1.0, 3 nearer beta row); tables must
    be created inside the context."""

    def __init__(self, variant):
        self.variant = int(variant)

    def __enter__(self):
        lib().orc_set_alg1_variant(self.variant)
        return self

    def __exit__(self, *exc):
        lib().orc_set_alg1_variant(0)
        return False


class freegas_procedure:
    """Context manager selecting the checker's free-gas procedure:
    "collision" (default; the device's two-body collision draw) or
    "reference" (restatement of NCFreeGasUtils.cc:492-930)."""

    def __init__(self, which):
        self.which = {"collision": 0, "reference": 1}[which]

    def __enter__(self):
        lib().orc_set_freegas_procedure(self.which)
        return self

    def __exit__(self, *exc):
        lib().orc_set_freegas_procedure(0)
        return False


def absoov(sigma_abs):
    return Proc(lib().orc_create_absoov(sigma_abs))


def powderbragg(v0_times_natoms, dsp, fsq_x_mult):
    d = np.ascontiguousarray(dsp, dtype=np.float64)
    f = np.ascontiguousarray(fsq_x_mult, dtype=np.float64)
    p = Proc(lib().orc_create_powderbragg(v0_times_natoms, _d(d), _d(f), d.size))
    return p


def pb_tables(p):
    ptr = p if isinstance(p, int) else p.ptr
    n = lib().orc_pb_ntables(ptr)
    e2d = np.empty(n)
    fdm = np.empty(n)
    lib().orc_pb_tables(ptr, _d(e2d), _d(fdm))
    return e2d, fdm


# process types of the checker's components
PROC_TYPES = {1: "PowderBragg", 2: "ElInc", 3: "SAB", 4: "Composition", 5: "SCBragg", 6: "FreeGas", 7: "AbsOOV"}


def components(p):
    """[(type name, scale, data)] of a composition after merging: data is
    (e2d, cumulative fdm) for PowderBragg, (msd, scaled bixs) for ElInc, the
    raw component pointer otherwise."""
    out = []
    for i in range(lib().orc_comp_ncomponents(p.ptr)):
        sc = np.zeros(1)
        t = lib().orc_comp_component(p.ptr, i, _d(sc))
        ptr = lib().orc_comp_component_ptr(p.ptr, i)
        if t == 1:
            data = pb_tables(ptr)
        elif t == 2:
            n = lib().orc_elinc_n(ptr)
            msd, bixs = np.empty(n), np.empty(n)
            lib().orc_elinc_data(ptr, _d(msd), _d(bixs))
            data = (msd, bixs)
        else:
            data = ptr
        out.append((PROC_TYPES[t], float(sc[0]), data))
    return out


def elinc(msd, bixs, scale):
    m = np.ascontiguousarray(msd, dtype=np.float64)
    b = np.ascontiguousarray(bixs, dtype=np.float64)
    s = np.ascontiguousarray(scale, dtype=np.float64)
    return Proc(lib().orc_create_elinc(m.size, _d(m), _d(b), _d(s)))


class SAB(Proc):
    def egrid(self):
        n = lib().orc_sab_negrid(self.ptr)
        eg = np.empty(n)
        xs = np.empty(n)
        lib().orc_sab_egrid(self.ptr, _d(eg), _d(xs))
        return eg, xs

    def analyse_xs(self, e):
        return lib().orc_sab_analyse_xs(self.ptr, e)

    def sampler(self, ie):
        ibl, nv, kind, fb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        lib().orc_sab_sampler_info(self.ptr, ie, ctypes.byref(ibl), ctypes.byref(nv), ctypes.byref(kind),
                                   ctypes.byref(fb))
        res = dict(ibeta_low=ibl.value, nvals=nv.value, kind=kind.value, firstbin=fb.value)
        if kind.value == 1:
            n = nv.value
            bx, by, bc = np.empty(n), np.empty(n), np.empty(n)
            idbl = np.empty(8 * max(n - 1, 1))
            iidx = np.empty(2 * max(n - 1, 1), dtype=np.int32)
            lib().orc_sab_sampler_arrays(self.ptr, ie, _d(bx), _d(by), _d(bc), _d(idbl), _i(iidx))
            res.update(bx=bx, by=by, bcdf=bc, info=idbl.reshape(-1, 8)[:n - 1], infoidx=iidx.reshape(-1, 2)[:n - 1])
        return res

    def derived(self, na, nb):
        ls = np.empty(na * nb)
        cu = np.empty(na * nb)
        lib().orc_sab_derived(self.ptr, _d(ls), _d(cu))
        return ls.reshape(nb, na), cu.reshape(nb, na)

    def sample_alpha(self, ie, ibeta, u):
        """the Alg1 alpha sampler of beta row ibeta at egrid point ie"""
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.empty(u.size)
        lib().orc_sab_sample_alpha(self.ptr, ie, ibeta, u.size, _d(u), _d(out))
        return out

    def sample_highe(self, ekin, seed=1, stream=0):
        """sampleHighE alone: (alpha, beta, nerr); alpha = -1 marks draws
        left to the table at Emax"""
        e = np.ascontiguousarray(np.atleast_1d(np.asarray(ekin, dtype=np.float64)))
        oa = np.empty(e.size)
        ob = np.empty(e.size)
        nerr = lib().orc_sab_sample_highe_many(self.ptr, seed, stream, _d(e), e.size, _d(oa), _d(ob))
        return oa, ob, nerr

    def sample_ab(self, ekin, seed=1, stream=0, nthreads=1):
        e = np.ascontiguousarray(np.atleast_1d(np.asarray(ekin, dtype=np.float64)))
        oa = np.empty(e.size)
        ob = np.empty(e.size)
        nerr = lib().orc_sab_sample_ab_many(self.ptr, seed, stream, _d(e), e.size, _d(oa), _d(ob), nthreads)
        return oa, ob, nerr


def sab(alpha, beta, sab_table, temperature, mass_amu, sigma_bound, suggested_emax=0.0, egrid=None):
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    s = np.ascontiguousarray(sab_table, dtype=np.float64).reshape(b.size, a.size)
    eg = np.ascontiguousarray(egrid if egrid is not None else np.zeros(0), dtype=np.float64)
    ptr = lib().orc_create_sab(a.size, b.size, _d(a), _d(b), _d(s), temperature, mass_amu, sigma_bound,
                               suggested_emax, eg.size, _d(eg) if eg.size else None)
    return SAB(ptr)


def composition(procs, scales):
    arr = (_vp * len(procs))(*[p.ptr for p in procs])
    sc = np.ascontiguousarray(scales, dtype=np.float64)
    p = Proc(lib().orc_create_composition(len(procs), arr, _d(sc)), keep=tuple(procs))
    return p


def hkl_cell_members(cell, species, dcutoff, fsquarecut=1e-5, merge_tolerance=1e-6):
    """enumeration without symmetry for any cell (a, b, c, alpha, beta, gamma)"""
    return hkl_orthogonal_members(*cell[:3], species, dcutoff, fsquarecut, merge_tolerance, angles=tuple(cell[3:]))


def hkl_orthogonal_members(a, b, c, species, dcutoff, fsquarecut=1e-5, merge_tolerance=1e-6, maxout=2000000,
                           maxmem=8000000, angles=None):
    natoms = np.array([len(s[2]) for s in species], dtype=np.int32)
    pos = np.ascontiguousarray(np.concatenate([np.asarray(s[2], dtype=np.float64).reshape(-1, 3)
                                               for s in species]).ravel())
    csl = np.array([s[0] for s in species], dtype=np.float64)
    msd = np.array([s[1] for s in species], dtype=np.float64)
    hkl = np.empty(3 * maxout, dtype=np.int32)
    mult = np.empty(maxout, dtype=np.int32)
    dsp = np.empty(maxout)
    fsq = np.empty(maxout)
    moff = np.empty(maxout + 1, dtype=np.int64)
    mhkl = np.empty(3 * maxmem, dtype=np.int32)
    if angles is None:
        n = lib().orc_hkl_orthogonal_members(a, b, c, len(species), _i(natoms), _d(pos), _d(csl), _d(msd), dcutoff,
                                             fsquarecut, merge_tolerance, maxout, _i(hkl), _i(mult), _d(dsp),
                                             _d(fsq), moff.ctypes.data_as(_i64p), _i(mhkl), maxmem)
    else:
        n = lib().orc_hkl_cell_members(a, b, c, *[float(x) for x in angles], len(species), _i(natoms), _d(pos),
                                       _d(csl), _d(msd), dcutoff, fsquarecut, merge_tolerance, maxout, _i(hkl),
                                       _i(mult), _d(dsp), _d(fsq), moff.ctypes.data_as(_i64p), _i(mhkl), maxmem)
    if n < 0:
        raise RuntimeError(f"oracle hkl enumeration failed ({n})")
    nm = int(moff[n])
    return (hkl[:3 * n].reshape(n, 3).copy(), mult[:n].copy(), dsp[:n].copy(), fsq[:n].copy(),
            moff[:n + 1].copy(), mhkl[:3 * nm].reshape(nm, 3).copy())


class SCB(Proc):
    def info(self):
        nf, nn, thr = ctypes.c_int(), ctypes.c_int64(), ctypes.c_double()
        gd = np.empty(15)
        lib().orc_scb_info(self.ptr, ctypes.byref(nf), ctypes.byref(nn), ctypes.byref(thr), _d(gd))
        return dict(nfam=nf.value, nnormals=nn.value, threshold=thr.value, gos=gd)

    def tables(self):
        inf = self.info()
        nf, nn, nlt = inf["nfam"], inf["nnormals"], int(inf["gos"][10])
        inv2d, xsf = np.empty(nf), np.empty(nf)
        off = np.empty(nf + 1, dtype=np.int64)
        nrm = np.empty((nn, 3))
        lt1, lt2 = np.empty(2 * nlt), np.empty(2 * nlt)
        lib().orc_scb_tables(self.ptr, _d(inv2d), _d(xsf), off.ctypes.data_as(_i64p), _d(nrm), _d(lt1), _d(lt2))
        return dict(inv2d=inv2d, xsfact=xsf, off=off, normals=nrm, lt_evalcosx=lt1.reshape(-1, 2),
                    lt_sofcosd=lt2.reshape(-1, 2))


def reci_lattice_orthogonal(a, b, c):
    """diag(2pi/a, 2pi/b, 2pi/c), row-major"""
    k = 6.283185307179586
    return np.array([k / a, 0.0, 0.0, 0.0, k / b, 0.0, 0.0, 0.0, k / c])


def scbragg(a, b, c, v0_times_natoms, dsp, fsq, mem_off, mem_hkl, cry2lab, mos_fwhm, prec=1e-3, ntrunc=0.0,
            reci=None):
    d = np.ascontiguousarray(dsp, dtype=np.float64)
    f = np.ascontiguousarray(fsq, dtype=np.float64)
    mo = np.ascontiguousarray(mem_off, dtype=np.int64)
    mh = np.ascontiguousarray(mem_hkl, dtype=np.int32).ravel()
    R = np.ascontiguousarray(cry2lab, dtype=np.float64).ravel()
    L = np.ascontiguousarray(reci_lattice_orthogonal(a, b, c) if reci is None else reci, dtype=np.float64).ravel()
    ptr = lib().orc_create_scbragg(_d(L), v0_times_natoms, d.size, _d(d), _d(f), mo.ctypes.data_as(_i64p),
                                   _i(mh), _d(R), mos_fwhm, prec, ntrunc)
    return SCB(ptr, keep=(L,))


def _lc_rotate(axis, v, c, s):
    """Rodrigues' rotation of unit vectors v (n, 3) by (cos, sin) (n,) around
    the unit axis, NCRotMatrix.hh:172-184: v c + (a x v) s + a ((a . v)(1 - c))"""
    cr = np.cross(axis, v)
    dt = v @ axis
    return v * c[:, None] + cr * s[:, None] + np.outer(dt * (1.0 - c), axis)


class LCBraggRef:
    """The layered-crystal reference models of
    ncrystal_core/src/extd_utils/NCLCRefModels.cc over a checker SCBragg
    process: mode > 0 LCBraggRef (cross section: mean over N equally spaced
    rotations phi_i = 2 pi i / N - pi around the layer axis, N the smallest
    prime >= mode, :25-52; scattering: `mode` random rotations, one picked by
    cross section, the crystal's scattering of the rotated neutron rotated
    back, :54-80), mode < 0 LCBraggRndmRot (-mode random rotations for both,
    :82-171).  Random numbers from numpy's PCG64."""

    def __init__(self, scb, lcaxis_lab, mode):
        self.scb = scb
        a = np.asarray(lcaxis_lab, dtype=np.float64)
        self.axis = a / np.linalg.norm(a)
        self.mode = int(mode)
        self.nsample = abs(self.mode)
        n = self.nsample
        while n < 2 or any(n % k == 0 for k in range(2, int(math.isqrt(n)) + 1)):
            n += 1
        self.nprime = n

    def _expand(self, ekin, dirs, c, s):
        """(n*N) energies and rotated directions; c, s of shape (n, N)"""
        n, N = c.shape
        u = dirs / np.linalg.norm(dirs, axis=1)[:, None]
        U = np.repeat(u, N, axis=0)
        return np.repeat(ekin, N), _lc_rotate(self.axis, U, c.ravel(), s.ravel())

    def xs_dir(self, ekin, dirs, seed=0, nthreads=8):
        ekin = np.asarray(ekin, dtype=np.float64)
        dirs = np.asarray(dirs, dtype=np.float64).reshape(-1, 3)
        n = len(ekin)
        if self.mode > 0:
            N = self.nprime
            phi = np.arange(N) * (2 * math.pi / N) - math.pi
            c, s = np.tile(np.cos(phi), (n, 1)), np.tile(np.sin(phi), (n, 1))
        else:
            N = self.nsample
            phi = np.random.default_rng(seed).uniform(-math.pi, math.pi, size=(n, N))
            c, s = np.cos(phi), np.sin(phi)
        E, D = self._expand(ekin, dirs, c, s)
        x = self.scb.xs_dir(E, np.ascontiguousarray(D), nthreads=nthreads).reshape(n, N)
        return np.array([math.fsum(r) for r in x]) / N

    def sample_dir(self, ekin, dirs, seed=0, nthreads=8):
        ekin = np.asarray(ekin, dtype=np.float64)
        dirs = np.asarray(dirs, dtype=np.float64).reshape(-1, 3)
        n, N = len(ekin), self.nsample
        rng = np.random.default_rng(seed)
        phi = rng.uniform(-math.pi, math.pi, size=(n, N))
        c, s = np.cos(phi), np.sin(phi)
        E, D = self._expand(ekin, dirs, c, s)
        x = self.scb.xs_dir(E, np.ascontiguousarray(D), nthreads=nthreads).reshape(n, N)
        cum = np.cumsum(x, axis=1)
        tot = cum[:, -1]
        live = tot > 0.0
        u = rng.random(n) * tot
        pick = np.minimum((cum <= u[:, None]).sum(axis=1), N - 1)
        sel = np.arange(n) * N + pick
        oE, oD, nerr = self.scb.sample_dir(E[sel], np.ascontiguousarray(D[sel]), seed=seed + 1, nthreads=nthreads)
        oD = np.asarray(oD).reshape(-1, 3)
        back = _lc_rotate(self.axis, oD, c.ravel()[sel], -s.ravel()[sel])
        out = np.where(live[:, None], back, dirs)
        return np.asarray(ekin).copy(), out, nerr


_LC_DISCR = 1099511627776.0  # 2^40


def _lc_discr(x):
    return int(x * _LC_DISCR + 0.5)


def _romberg(f_many, a, b, acc):
    """NCRomberg.cc:58-140 (the library's romberg restatement): 17 points,
    then halvings up to level 15, accepting |R_k - R_{k-1}| <= acc |R_k|;
    f_many(x array) -> values"""
    h = b - a
    fv = f_many(a + (h / 16.0) * np.arange(17))
    h *= 0.5
    R = [[(fv[0] + fv[16]) * h]]
    R.append([h * fv[8] + 0.5 * R[0][0]])
    h *= 0.5
    R.append([h * (fv[4] + fv[12]) + 0.5 * R[1][0]])
    h *= 0.5
    R.append([h * ((fv[2] + fv[6]) + (fv[10] + fv[14])) + 0.5 * R[2][0]])
    h *= 0.5
    R.append([h * (((fv[1] + fv[3]) + (fv[5] + fv[7])) + ((fv[9] + fv[11]) + (fv[13] + fv[15]))) + 0.5 * R[3][0]])
    for i in range(1, 5):
        nk = 1.0
        for j in range(i):
            nk *= 4.0
            R[i].append((nk * R[i][j] - R[i - 1][j]) / (nk - 1.0))
    if abs(R[4][4] - R[3][3]) <= acc * abs(R[4][4]):
        return R[4][4]
    prev, nj = R[4], 8
    for i in range(5, 16):
        hh = h
        h *= 0.5
        nj *= 2
        c = float(np.sum(f_many(a + h + hh * np.arange(nj))))
        row = [h * c + 0.5 * prev[0]]
        nk = 1.0
        for j in range(i):
            nk *= 4.0
            row.append((nk * row[j] - prev[j]) / (nk - 1.0))
        if abs(row[i] - prev[i - 1]) <= acc * abs(row[i]):
            return row[i]
        prev = row
    return prev[-1]


class LCHelperXS:
    """Cross sections of NCrystal's layered-crystal model (lcmode=0,
    ncrystal_core/src/extd_utils/NCLCUtils.cc): plane sets of equal
    (d, angle to the layer axis) merged on a 2^-40 grid and sorted by
    decreasing (d, angle) (:58-131); per neutron the discretised wavelength
    and |cos| of its angle to the axis (:332-344), the phi ranges of
    crystallite rotations that reach the mosaic band of each plane set's
    normal and anti-normal (LCROIFinder, :159-303), the Romberg average of the
    Gaussian-mosaic cross section over each range (:396-439, 455-516), on-axis
    special cases, times 1/(V natoms).  scb: the checker's SCBragg of the
    crystal (its mosaic); normals: crystal-frame demi-normals with their
    family's |F|^2 and d."""

    def __init__(self, scb, normals, fsq, dsp, lcaxis_crys, lcaxis_lab, v0n):
        self.scb = scb
        nf, nn, thr, gd = ctypes.c_int(), ctypes.c_int64(), ctypes.c_double(), np.empty(15)
        lib().orc_scb_info(scb.ptr, ctypes.byref(nf), ctypes.byref(nn), ctypes.byref(thr), _d(gd))
        self.trunc, self.prec, self.cta, self.sta = gd[1], gd[2], gd[3], gd[4]
        ax = np.asarray(lcaxis_crys, dtype=np.float64)
        ax = ax / np.linalg.norm(ax)
        lab = np.asarray(lcaxis_lab, dtype=np.float64)
        self.axis = lab / np.linalg.norm(lab)
        sets = {}
        for n, f, d in zip(np.asarray(normals).reshape(-1, 3), fsq, dsp):
            ca = abs(float(ax @ n))
            alpha = 0.0 if ca > 0.9999999 else math.acos(ca)
            key = (_lc_discr(d), _lc_discr(alpha))
            sets[key] = sets.get(key, 0.0) + f
        self.sets = []
        for (kd, ka), f in sorted(sets.items(), reverse=True):
            d, alpha = kd / _LC_DISCR, min(ka / _LC_DISCR, math.pi / 2)
            self.sets.append(dict(twodsp=2 * d, inv2d=0.5 / d, c1=math.cos(alpha) if alpha else 1.0,
                                  s1=math.sin(alpha) if alpha else 0.0,
                                  c1minus=math.cos(alpha - self.trunc) if alpha > self.trunc else 1.0,
                                  c1plus=math.cos(alpha + self.trunc), fsq=f))
        self.xsfact = 1.0 / v0n

    def _raw(self, wl, ps, cosg):
        cosg = np.ascontiguousarray(np.atleast_1d(cosg), dtype=np.float64)
        out = np.empty(cosg.size)
        lib().orc_scb_rawxs(self.scb.ptr, wl, ps["inv2d"], ps["fsq"], cosg.size, _d(cosg), _d(out))
        return out

    def _rois(self, wl, c3, s3, ps):
        c2 = ps["inv2d"] * wl
        c23 = c3 * c2
        s2 = math.sqrt(1.0 - c2 * c2)
        s23 = s3 * s2
        lo, hi = c23 - s23, c23 + s23
        if max(lo, ps["c1plus"]) > min(hi, ps["c1minus"]):
            return []
        anti = max(lo, -ps["c1minus"]) <= min(hi, -ps["c1plus"])
        signs = [1.0] + ([-1.0] if anti else [])
        if not ps["s1"]:
            return [("axis", sg, 0.0, 0.0) for sg in signs]
        if abs(s3) < 1e-10:
            return [("neutron", sg, 0.0, 0.0) for sg in signs]
        c2ta, s2ta = c2 * self.cta, s2 * self.sta
        c2low = c2ta - s2ta
        c2high = 1.0 if s2 < self.sta else c2ta + s2ta
        a = ps["c1"] * c3
        b = 1.0 / (ps["s1"] * s3)
        mab, lowb, highb = -a * b, c2low * b, c2high * b
        out = []
        for sg in signs:
            p1 = max(-1.0, min(1.0, mab + sg * lowb))
            p2 = max(-1.0, min(1.0, mab + sg * highb))
            if abs(p1 - p2) > 1e-10:
                out.append(("range", sg, math.acos(max(p1, p2)), math.acos(min(p1, p2))))
        return out

    def xs_one(self, wl_raw, u):
        wl = _lc_discr(wl_raw) / _LC_DISCR
        c3 = min(_lc_discr(abs(float(self.axis @ u))) / _LC_DISCR, 1.0)
        s3 = math.sqrt(abs(1.0 - c3 * c3))
        acc = min(max(self.prec, 1e-7), 1e-2)
        tot = 0.0
        for ps in self.sets:
            if wl > ps["twodsp"]:
                break
            for kind, sg, r0, r1 in self._rois(wl, c3, s3, ps):
                if kind == "axis":
                    x = self._raw(wl, ps, sg * c3)[0]
                elif kind == "neutron":
                    x = self._raw(wl, ps, sg * ps["c1"] * c3)[0]
                else:
                    k1, k0 = ps["s1"] * s3 * sg, ps["c1"] * sg * c3
                    x = _romberg(lambda phi: self._raw(wl, ps, k1 * np.cos(phi) + k0), r0, r1, acc) / math.pi
                tot += x
        return self.xsfact * tot

    def xs_dir(self, ekin, dirs):
        ekin = np.atleast_1d(np.asarray(ekin, dtype=np.float64))
        dirs = np.asarray(dirs, dtype=np.float64).reshape(-1, 3)
        out = np.empty(len(ekin))
        for i, (e, d) in enumerate(zip(ekin, dirs)):
            out[i] = self.xs_one(math.sqrt(0.081804209605330899 / e), d / np.linalg.norm(d)) if e > 0 else 0.0
        return out


def debug_sort_fams(dsp, fsq, mult, hkl):
    d = np.ascontiguousarray(dsp, dtype=np.float64)
    f = np.ascontiguousarray(fsq, dtype=np.float64)
    m = np.ascontiguousarray(mult, dtype=np.int32)
    h = np.ascontiguousarray(hkl, dtype=np.int32).ravel()
    perm = np.empty(d.size, dtype=np.int32)
    lib().orc_debug_sort_fams(d.size, _d(d), _d(f), _i(m), _i(h), _i(perm))
    return perm
