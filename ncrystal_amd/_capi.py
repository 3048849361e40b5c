"""ctypes bindings of libncrystal_amd.so (C ABI in include/ncrystal_amd.h).

Mirrors how NCrystal's Python layer binds its C API
(ncrystal_python/src/NCrystal/_chooks.py:65-78 load, :134-191 error
handling): the library is told not to halt or print on errors, and every
call is followed by an error check that raises a Python exception.
The library is always the in-tree build; a missing library is an error
(there is no CPU substitute).
"""
import ctypes
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIBPATH = _HERE / "libncrystal_amd.so"


class NCAMDError(RuntimeError):
    pass


class NCBadInput(NCAMDError):
    pass


class NCCalcError(NCAMDError):
    pass


class NCLogicError(NCAMDError):
    pass


class NCNoDevice(NCAMDError):
    pass


_ERRMAP = {"BadInput": NCBadInput, "CalcError": NCCalcError, "LogicError": NCLogicError,
           "NoDevice": NCNoDevice}

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_ul = ctypes.c_ulong
_vp = ctypes.c_void_p


class Handle(ctypes.Structure):
    _fields_ = [("internal", ctypes.c_void_p)]


class BatchHandle(ctypes.Structure):
    _fields_ = [("internal", ctypes.c_void_p)]


class MlpHandle(ctypes.Structure):
    _fields_ = [("internal", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIBPATH.exists():
        raise NCAMDError(f"{LIBPATH} is not built (run ncrystal_amd.build.build())")
    L = ctypes.CDLL(str(LIBPATH))
    sig = {
        "ncamd_version_str": (ctypes.c_char_p, []),
        "ncamd_device_count": (ctypes.c_int, []),
        "ncamd_error": (ctypes.c_int, []),
        "ncamd_lasterror": (ctypes.c_char_p, []),
        "ncamd_lasterrortype": (ctypes.c_char_p, []),
        "ncamd_clearerror": (None, []),
        "ncamd_sethaltonerror": (ctypes.c_int, [ctypes.c_int]),
        "ncamd_setquietonerror": (ctypes.c_int, [ctypes.c_int]),
        "ncamd_clone": (Handle, [Handle, ctypes.c_uint64]),
        "ncamd_sab_std_format": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                                _ip, _ip, _dp, _dp, _dp, _i64p]),
        "ncamd_vdos_regularise": (ctypes.c_int, [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]),
        "ncamd_vdos_analyse": (ctypes.c_int, [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
                                              _dp]),
        "ncamd_vdos_gn": (ctypes.c_int, [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]),
        "ncamd_vdos_kernel": (ctypes.c_int, [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int, ctypes.c_double, _ip, _ip, _dp, _dp, _dp, _dp]),
        "ncamd_create_sabscatter_from_vdos": (Handle, [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_double,
                                                       ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                       ctypes.c_double]),
        "ncamd_ref": (None, [Handle]),
        "ncamd_unref": (None, [ctypes.POINTER(Handle)]),
        "ncamd_refcount": (ctypes.c_int, [Handle]),
        "ncamd_valid": (ctypes.c_int, [Handle]),
        "ncamd_isoriented": (ctypes.c_int, [Handle]),
        "ncamd_debye_msd": (ctypes.c_double, [ctypes.c_double] * 3),
        "ncamd_hkl_orthogonal": (ctypes.c_int64, [ctypes.c_double] * 3 + [ctypes.c_int, _ip, _dp, _dp, _dp,
                                                                         ctypes.c_double, ctypes.c_double,
                                                                         ctypes.c_double, ctypes.c_int64, _ip,
                                                                         _ip, _dp, _dp, _i64p]),
        "ncamd_hkl_orthogonal_members": (ctypes.c_int64, [ctypes.c_double] * 3 + [ctypes.c_int, _ip, _dp, _dp, _dp,
                                                                                 ctypes.c_double, ctypes.c_double,
                                                                                 ctypes.c_double, ctypes.c_int64,
                                                                                 _ip, _ip, _dp, _dp, _i64p,
                                                                                 ctypes.c_int64, _i64p, _ip, _i64p]),
        "ncamd_hkl_cell_members": (ctypes.c_int64, [ctypes.c_double] * 6 + [ctypes.c_int, _ip, _dp, _dp, _dp,
                                                                                 ctypes.c_double, ctypes.c_double,
                                                                                 ctypes.c_double, ctypes.c_int64,
                                                                                 _ip, _ip, _dp, _dp, _i64p,
                                                                                 ctypes.c_int64, _i64p, _ip, _i64p]),
        "ncamd_hkl_spacegroup_members": (ctypes.c_int64, [ctypes.c_int] + [ctypes.c_double] * 6 +
                                         [ctypes.c_int, _ip, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int64, _ip, _ip, _dp, _dp, _i64p, ctypes.c_int64, _i64p, _ip,
                                          _i64p]),
        "ncamd_reci_lattice": (ctypes.c_int, [ctypes.c_double] * 6 + [_dp]),
        "ncamd_create_freegas": (Handle, [ctypes.c_double] * 3),
        "ncamd_create_absoov": (Handle, [ctypes.c_double]),
        "ncamd_create_scbragg": (Handle, [ctypes.c_double, ctypes.c_int64, _dp, _dp, _i64p, _ip, _dp, _dp,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "ncamd_create_lcbragg_ref": (Handle, [Handle, _dp, ctypes.c_int]),
        "ncamd_create_lcbragg": (Handle, [ctypes.c_double, ctypes.c_int64, _dp, _dp, _i64p, _ip, _dp, _dp, _dp,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
        "ncamd_scbragg_info": (ctypes.c_int, [Handle, _i64p, _i64p, _dp, _dp]),
        "ncamd_scbragg_tables": (ctypes.c_int, [Handle, _dp, _dp, _i64p, _dp, _dp, _dp]),
        "ncamd_create_powderbragg": (Handle, [ctypes.c_double, _dp, _dp, ctypes.c_int64, ctypes.c_int]),
        "ncamd_create_elincscatter": (Handle, [ctypes.c_int, _dp, _dp, _dp]),
        "ncamd_create_sabscatter": (Handle, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                             _dp]),
        "ncamd_create_composition": (Handle, [ctypes.c_int, ctypes.POINTER(Handle), _dp]),
        "ncamd_composition_component": (ctypes.c_int, [Handle, ctypes.c_int, ctypes.POINTER(Handle), _dp]),
        "ncamd_powderbragg_tables": (ctypes.c_int64, [Handle, _dp, _dp, _fp, ctypes.c_int64]),
        "ncamd_sab_egrid": (ctypes.c_int, [Handle, _dp, _dp, ctypes.c_int]),
        "ncamd_sab_derived": (ctypes.c_int, [Handle, _dp, _dp]),
        "ncamd_sab_sampler": (ctypes.c_int, [Handle, ctypes.c_int, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _ip]),
        "ncamd_setrngseed": (None, [Handle, ctypes.c_uint64]),
        "ncamd_getrngstate": (None, [Handle, _u64p, _u64p]),
        "ncamd_setrngstate": (None, [Handle, ctypes.c_uint64, ctypes.c_uint64]),
        "ncamd_crosssection_nonoriented_many": (None, [Handle, _dp, _ul, _ul, _dp]),
        "ncamd_crosssection_nonoriented_many_fp32": (None, [Handle, _dp, _ul, _ul, _dp]),
        "ncamd_samplescatterisotropic_many": (None, [Handle, _dp, _ul, _ul, _dp, _dp]),
        "ncamd_samplescatter_many": (None, [Handle, ctypes.c_double, ctypes.POINTER(ctypes.c_double * 3), _ul,
                                            _dp, _dp, _dp, _dp]),
        "ncamd_crosssection_many_dirs": (None, [Handle, _dp, _dp, _ul, _dp]),
        "ncamd_samplescatter_many_dirs": (None, [Handle, _dp, _dp, _ul, _dp, _dp]),
        "ncamd_dev_crosssection_nonoriented_many": (None, [Handle, _vp, _ul, _ul, _vp, _vp]),
        "ncamd_dev_crosssection_nonoriented_many_fp32": (None, [Handle, _vp, _ul, _ul, _vp, _vp]),
        "ncamd_dev_samplescatterisotropic_many": (None, [Handle, _vp, _ul, _ul, _vp, _vp, _vp]),
        "ncamd_dev_samplescatter_many_dirs": (None, [Handle, _vp, _vp, _ul, _vp, _vp, _vp]),
        "ncamd_dev_crosssection_many_dirs": (None, [Handle, _vp, _vp, _ul, _vp, _vp]),
        "ncamd_dev_sab_sample_alphabeta": (None, [Handle, _vp, _ul, _vp, _vp, _vp]),
        "ncamd_dev_check_errors": (ctypes.c_int64, [Handle, _vp]),
        "ncamd_sab_use_envelope_sampler": (ctypes.c_int, [Handle, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                          ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "ncamd_sab_use_plain_sampler": (ctypes.c_int, [Handle]),
        "ncamd_sab_envelope_tables": (ctypes.c_int, [Handle, _dp, _dp]),
        "ncamd_sab_envelope_stats": (ctypes.c_int, [Handle, _dp]),
        "ncamd_sab_envelope_squeeze": (ctypes.c_int, [Handle, ctypes.c_int]),
        "ncamd_dev_sab_sample_alphabeta_envelope": (ctypes.c_int64, [Handle, _vp, _ul, _vp, _vp, _vp]),
        "ncamd_mlp_create": (MlpHandle, [ctypes.c_uint64, ctypes.c_double, ctypes.c_double]),
        "ncamd_mlp_destroy": (None, [ctypes.POINTER(MlpHandle)]),
        "ncamd_mlp_nparams": (ctypes.c_int, []),
        "ncamd_mlp_get_params": (ctypes.c_int, [MlpHandle, _fp]),
        "ncamd_mlp_set_params": (ctypes.c_int, [MlpHandle, _fp]),
        "ncamd_mlp_forward": (ctypes.c_int, [MlpHandle, _dp, ctypes.c_int, _fp]),
        "ncamd_mlp_loss_grad": (ctypes.c_double, [MlpHandle, _dp, _dp, _dp, ctypes.c_int, _fp]),
        "ncamd_minimc_sphere": (ctypes.c_int, [Handle, Handle, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp,
                                               ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_int, _dp, ctypes.c_uint64, _dp, _dp]),
        "ncamd_mlp_train_step": (ctypes.c_double, [MlpHandle, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double]),
        "ncamd_mlp_train_sab": (ctypes.c_double, [MlpHandle, Handle, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                  ctypes.c_uint64]),
        "ncamd_debug_gemm_bf16": (ctypes.c_int, [_fp, _fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "ncamd_sab_use_learned_sampler": (ctypes.c_int, [Handle, MlpHandle, ctypes.c_int, ctypes.c_double,
                                                         ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "ncamd_batch_create": (BatchHandle, [ctypes.c_int, ctypes.POINTER(Handle), _i64p]),
        "ncamd_batch_destroy": (None, [ctypes.POINTER(BatchHandle)]),
        "ncamd_batch_total": (ctypes.c_int64, [BatchHandle]),
        "ncamd_dev_batch_crosssection": (None, [BatchHandle, _vp, _vp, _vp]),
        "ncamd_dev_batch_samplescatterisotropic": (None, [BatchHandle, _vp, _vp, _vp, _vp]),
        "ncamd_dev_batch_check_errors": (ctypes.c_int64, [BatchHandle, _vp]),
        "ncamd_batch_get_tuning": (None, [BatchHandle, _ip, _ip, _ip]),
        "ncamd_batch_set_tuning": (None, [BatchHandle, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "ncamd_batch_graph_stats": (None, [BatchHandle, ctypes.c_int, _i64p, _i64p]),
        "ncamd_dev_batch_autotune": (ctypes.c_double, [BatchHandle, _vp, _vp, _vp, _vp]),
        "ncamd_dev_malloc": (_vp, [ctypes.c_uint64]),
        "ncamd_dev_free": (None, [_vp]),
        "ncamd_memcpy_htod": (None, [_vp, _vp, ctypes.c_uint64]),
        "ncamd_memcpy_dtoh": (None, [_vp, _vp, ctypes.c_uint64]),
        "ncamd_memcpy_dtod": (None, [_vp, _vp, ctypes.c_uint64, _vp]),
        "ncamd_stream_create": (_vp, []),
        "ncamd_stream_destroy": (None, [_vp]),
        "ncamd_stream_sync": (None, [_vp]),
        "ncamd_event_create": (_vp, []),
        "ncamd_event_destroy": (None, [_vp]),
        "ncamd_event_record": (None, [_vp, _vp]),
        "ncamd_event_elapsed_ms": (ctypes.c_float, [_vp, _vp]),
        "ncamd_dev_fill_loguniform": (None, [_vp, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_uint64, _vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L.ncamd_sethaltonerror(0)
    L.ncamd_setquietonerror(1)
    _lib = L
    return L


def check():
    L = lib()
    if L.ncamd_error():
        msg = L.ncamd_lasterror().decode()
        typ = L.ncamd_lasterrortype().decode()
        L.ncamd_clearerror()
        raise _ERRMAP.get(typ, NCAMDError)(f"[{typ}] {msg}")


def call(name, *args):
    r = getattr(lib(), name)(*args)
    check()
    return r


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def i64ptr(a):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_i64p)


def fptr(a):
    return a.ctypes.data_as(_fp)


def as_f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


class DeviceArray:
    """A float64 device buffer owned by libncrystal_amd (hipMalloc/hipFree)."""

    def __init__(self, n):
        self.n = int(n)
        self.ptr = call("ncamd_dev_malloc", 8 * max(self.n, 1))

    @classmethod
    def from_numpy(cls, a):
        a = as_f64(a)
        d = cls(a.size)
        call("ncamd_memcpy_htod", d.ptr, a.ctypes.data, 8 * a.size)
        return d

    def numpy(self, count=None):
        count = self.n if count is None else int(count)
        out = np.empty(count)
        call("ncamd_memcpy_dtoh", out.ctypes.data, self.ptr, 8 * count)
        return out

    def free(self):
        if self.ptr:
            lib().ncamd_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.ptr = call("ncamd_stream_create")

    def sync(self):
        call("ncamd_stream_sync", self.ptr)

    def __del__(self):
        try:
            lib().ncamd_stream_destroy(self.ptr)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.ptr = call("ncamd_event_create")

    def record(self, stream):
        call("ncamd_event_record", self.ptr, stream.ptr if isinstance(stream, Stream) else stream)

    def elapsed_ms(self, end):
        return call("ncamd_event_elapsed_ms", self.ptr, end.ptr)

    def __del__(self):
        try:
            lib().ncamd_event_destroy(self.ptr)
        except Exception:
            pass
