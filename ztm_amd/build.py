"""Build libztm_gfm.so (HIP, gfx950) and the GFM program in-tree.

    python -m ztm_amd.build

Every HIP source is compiled for gfx950 only, with packed FP32 instructions
disabled (-packed-fp32-ops), into position-independent objects that are linked
into ztm_amd/libztm_gfm.so. The GFM program (ztm_amd/bin/GFM) is plain C++
linked against that library with an $ORIGIN rpath, so the pair can be copied
next to a TFM build as one unit.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libztm_gfm.so")
GFM = os.path.join(PKG, "bin", "GFM")

HIP_SOURCES = ["gfm_abi.hip", "gfm_batch.hip", "gfm_batched.hip", "gfm_batched64.hip",
               "gfm_calib.hip", "gfm_gn.hip", "gfm_noalloc.hip",
               "gfm_pointmass.hip", "gfm_prism.hip", "gfm_reg.hip", "gfm_stored.hip",
               "gfm_tfm.hip"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unknown-pragmas",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def torch_ext_path() -> str:
    import sysconfig
    return os.path.join(PKG, "_ztm_torch" + sysconfig.get_config_var("EXT_SUFFIX"))


def _build_torch_ext() -> str:
    """libtorch autograd bindings: plain g++ against torch's headers and
    libraries (the HIP stream comes from c10/hip), linked to libztm_gfm."""
    import sysconfig
    import torch
    tdir = os.path.dirname(torch.__file__)
    tinc, tlib = os.path.join(tdir, "include"), os.path.join(tdir, "lib")
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    out = torch_ext_path()
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-w",
          os.path.join(CSRC, "gfm_torch.cpp"), "-o", out + ".tmp",
          "-I" + tinc, "-I" + os.path.join(tinc, "torch", "csrc", "api", "include"),
          "-I" + sysconfig.get_paths()["include"], "-I" + os.path.join(ROCM, "include"),
          "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1", "-DTORCH_EXTENSION_NAME=_ztm_torch",
          "-DTORCH_API_INCLUDE_EXTENSION_H", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
          "-L" + tlib, "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_python",
          "-L" + PKG, "-lztm_gfm", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + tlib])
    os.replace(out + ".tmp", out)
    return out


def build(verbose: bool = True, torch_ext: bool = True) -> str:
    """Compile every HIP source for gfx950 and link the library, the GFM
    program and the libtorch bindings."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(GFM), exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(HIP_SOURCES))) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", GFM, os.path.join(CSRC, "gfm_main.cpp"),
          "-L" + PKG, "-lztm_gfm", "-Wl,-rpath,$ORIGIN/..",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    built = [LIB, GFM]
    if torch_ext:
        built.append(_build_torch_ext())
    if verbose:
        print("built " + ", ".join(os.path.relpath(p, ROOT) for p in built))
    return LIB


if __name__ == "__main__":
    build()
