"""C-ABI library: loads without a GPU, exports every declared symbol, and
fails loudly (no CPU fallback) when no device is present.  CPU only."""
import ctypes
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_symbols():
    txt = (ROOT / "include" / "ncrystal_amd.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ncamd_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("ncamd_crosssection_nonoriented_many", "ncamd_samplescatterisotropic_many",
              "ncamd_samplescatter_many", "ncamd_create_sabscatter", "ncamd_hkl_orthogonal"):
        assert s in syms


def test_library_exports_all_declared_symbols():
    from ncrystal_amd import _capi
    L = ctypes.CDLL(str(_capi.LIBPATH))
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_error_convention_invalid_handle():
    from ncrystal_amd import _capi as C
    L = C.lib()
    h = C.Handle()
    h.internal = None
    out = np.zeros(3)
    e = np.ones(3)
    L.ncamd_crosssection_nonoriented_many(h, C.dptr(e), 3, 1, C.dptr(out))
    # poisoned output and an error raised through the error state
    assert np.all(out == -1.0)
    with pytest.raises(C.NCAMDError):
        C.check()


def test_no_device_fails_loudly():
    from ncrystal_amd import _capi as C
    L = C.lib()
    if L.ncamd_device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(C.NCNoDevice):
        from ncrystal_amd import core
        core.create_powderbragg(10.0, [1.0, 2.0], [1.0, 1.0])
