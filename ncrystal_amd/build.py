"""Build libncrystal_amd.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m ncrystal_amd.build   (or call build()).
"""
import os
import pathlib
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libncrystal_amd.so"
BUILD = HERE / "_build"
SOURCES = ["nc_kernels.hip", "nc_sab_build.hip", "nc_hkl.hip", "nc_scb.hip", "nc_batch.hip", "nc_kb.hip", "nc_mlp.hip", "nc_vdos.hip", "nc_mmc.hip", "nc_lc.hip", "nc_capi.hip"]
ARCH = os.environ.get("NCAMD_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # the reference-parity path relies on the written operation order
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = BUILD / (src.replace(".hip", ".o"))
    cmd = [hipcc(), "-c", str(CSRC / src), "-o", str(obj)] + HIPCC_FLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    BUILD.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-o", str(tmp)] + [str(o) for o in objs] + [f"--offload-arch={ARCH}", "-fPIC"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
