"""In-tree build of the native pieces.

* ``libpolysplat_amd.so``: the HIP kernels and the C ABI of
  ``include/polysplat_amd.h``, compiled by hipcc for gfx950.
* ``_C*.so``: the pybind11 module that mirrors the reference's
  ``polysplat.ops`` surface (csrc/pybind.cpp:1154-1372) on torch tensors and
  calls the C ABI.

Both land next to this file so they travel with the source tree.  Run
``python polysplat_amd/_build.py`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_objs"
LIB = PKG / "libpolysplat_amd.so"
ARCH = os.environ.get("POLYSPLAT_ARCH", "gfx950")

HIP_SOURCES = {
    # source: extra flags
    "preprocess.hip": ["-ffp-contract=off"],
    "sort.hip": [],
    "ranges_gather.hip": [],
    "render.hip": [],
    "forward.hip": [],
    "tile_order.hip": [],
}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)


def _compile_hip(src: str, extra: list[str]) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    _run([HIPCC, *COMMON, *extra, "-c", str(CSRC / src), "-o", str(obj)])
    return obj


def build_native(jobs: int = 4) -> Path:
    BUILD.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda kv: _compile_hip(*kv), HIP_SOURCES.items()))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


def _torch_flags() -> tuple[list[str], list[str]]:
    import torch
    from torch.utils import cpp_extension

    inc = [f"-I{p}" for p in cpp_extension.include_paths()]
    inc += [f"-I{sysconfig.get_paths()['include']}", "-I/opt/rocm/include"]
    import pybind11

    inc.append(f"-I{pybind11.get_include()}")
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    defs = [
        f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
        "-DTORCH_API_INCLUDE_EXTENSION_H",
        "-DTORCH_EXTENSION_NAME=_C",
        "-D__HIP_PLATFORM_AMD__=1",
        "-DUSE_ROCM=1",
    ]
    libdir = Path(torch.__file__).resolve().parent / "lib"
    libs = [
        f"-L{libdir}", f"-Wl,-rpath,{libdir}",
        "-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python", "-lc10_hip", "-ltorch_hip",
        "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lamdhip64",
        f"-L{PKG}", "-Wl,-rpath,$ORIGIN", "-lpolysplat_amd",
    ]
    return inc + defs, libs


def build_pybind() -> Path:
    cflags, libs = _torch_flags()
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = PKG / f"_C{suffix}"
    tmp = out.with_name(out.name + ".tmp")
    BUILD.mkdir(exist_ok=True)

    def cc(src: str) -> Path:
        obj = BUILD / (Path(src).stem + ".o")
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-variable",
              "-ffp-contract=off", *cflags, f"-I{ROOT / 'include'}", "-c", str(CSRC / src), "-o",
              str(obj)])
        return obj

    with ThreadPoolExecutor(max_workers=2) as ex:
        objs = list(ex.map(cc, ["ops_bind.cpp", "ply.cpp"]))
    _run(["g++", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp), *libs])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_native()
    build_pybind()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("native", "all"):
        build_native()
    if what in ("pybind", "all"):
        build_pybind()
