"""Builds the native parts of strns2cards_amd in-tree.

* ``libstrns2cards_hip.so`` — the gfx950 HIP kernels behind the C ABI in
  ``include/strns2cards_hip.h`` (hipcc, one object per ``csrc/*.hip``).
* ``_core*.so`` — the PyTorch-ROCm pybind11 binding of that ABI (g++).

Run ``python strns2cards_amd/build.py`` (or ``__graft_entry__.build()``); it is
loaded by path because the package itself does not import before these
libraries exist.
"""
from __future__ import annotations

import glob
import hashlib
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libstrns2cards_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["scan.hip", "binning.hip", "dda.hip", "dda_mask.hip", "raster.hip", "bitset.hip",
               "antialias.hip", "curve.hip", "fit.hip", "coverage.hip", "gridsolve.hip",
               "cardmesh.hip", "texture.hip", "adam.hip"]

# Contraction off and correctly rounded divide/sqrt are the numerics contract
# with the CPU oracle; packed fp32 is disabled for the device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{os.path.join(ROOT, 'include')}",
]


def source_hash() -> str:
    """Hash of what the libraries are built from: every file of csrc/ and
    include/ and the compiler flags (16 hex digits), independent of where
    the tree lies."""
    h = hashlib.sha256(" ".join(HIP_FLAGS).replace(ROOT, "<root>").encode())
    files = sorted(glob.glob(os.path.join(CSRC, "*")) + glob.glob(os.path.join(ROOT, "include", "*")))
    for f in files:
        if os.path.isfile(f):
            h.update(os.path.relpath(f, ROOT).encode())
            with open(f, "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()[:16]


def _id_flag() -> str:
    return f'-DS2C_BUILD_ID="{source_hash()}"'


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + proc.stdout)
    noisy = [l for l in proc.stdout.splitlines()
             if l.strip() and "packed-fp32-ops" not in l and "is not a recognized feature" not in l]
    if noisy:
        sys.stderr.write("\n".join(noisy) + "\n")


def _compile_hip(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, _id_flag(), "-c", os.path.join(CSRC, src), "-o", out])
    return out


def build_hip(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(HIP_SOURCES)))) as ex:
        objs = list(ex.map(_compile_hip, HIP_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_core() -> str:
    import torch.utils.cpp_extension as ce
    import torch

    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "_core" + suffix)
    incs = []
    for p in ce.include_paths(device_type="cuda"):
        incs += ["-I", p]
    libs = []
    for p in ce.library_paths(device_type="cuda"):
        libs += ["-L", p]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = [
        "g++", "-O2", "-std=c++17", "-fPIC", "-shared",
        os.path.join(CSRC, "core_binding.cpp"), "-o", out + ".tmp",
        *incs, "-I", sysconfig.get_paths()["include"],
        "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
        "-DTORCH_EXTENSION_NAME=_core", "-DTORCH_API_INCLUDE_EXTENSION_H", _id_flag(),
        f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
        *libs, "-L", HERE,
        "-lstrns2cards_hip", "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu",
        "-ltorch_python", "-ltorch_hip",
        "-Wl,-rpath,$ORIGIN",
    ]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_all() -> None:
    build_hip()
    build_core()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
