"""The NCrystal C API names (include/ncrystal_dropin.h): a C program compiled
against the header and linked with the library runs its device-free checks
(tests/c/test_dropin_abi.c), and the library exports every symbol the two
headers declare.  CPU only."""
import ctypes
import re
import shutil
import subprocess

import pytest

from conftest import ROOT


def _declared(header, prefix):
    txt = (ROOT / "include" / header).read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", txt)))


def test_dropin_header_declares_the_reference_names():
    syms = _declared("ncrystal_dropin.h", "ncrystal_")
    for s in ("ncrystal_crosssection_nonoriented_many", "ncrystal_samplescatterisotropic_many",
              "ncrystal_samplescatter_many", "ncrystal_isnonoriented", "ncrystal_domain", "ncrystal_version_str",
              "ncrystal_unref", "ncrystal_clone_scatter", "ncrystal_samplescatter_rs"):
        assert s in syms


def test_library_exports_dropin_symbols():
    from ncrystal_amd import _capi
    L = ctypes.CDLL(str(_capi.LIBPATH))
    syms = _declared("ncrystal_dropin.h", "ncrystal_") + _declared("ncrystal_dropin.h", "ncamd_")
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_c_program_against_header(tmp_path):
    from ncrystal_amd import _capi
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    exe = tmp_path / "test_dropin_abi"
    lib = _capi.LIBPATH
    r = subprocess.run([cc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", str(ROOT / "include"),
                        str(ROOT / "tests" / "c" / "test_dropin_abi.c"), "-o", str(exe), str(lib),
                        f"-Wl,-rpath,{lib.parent}", "-lm"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
