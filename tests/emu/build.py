"""Builds the host emulation of the HIP kernels (test infrastructure only).

The unmodified kernel sources under polysplat_amd/csrc are compiled as C++
for the CPU against the HIP shim in tests/emu/include, together with the
work-item scheduler, into tests/emu/_build/libpolysplat_emu.so, which
exports the same C ABI as the GPU library.  With --asan it is built under
AddressSanitizer (load it with libasan preloaded).

    python tests/emu/build.py [--asan]
"""
from __future__ import annotations

import argparse
import fcntl
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
CSRC = ROOT / "polysplat_amd" / "csrc"
SOURCES = ["preprocess.hip", "sort.hip", "ranges_gather.hip", "render.hip", "forward.hip",
           "tile_order.hip"]


def build(asan: bool = False) -> Path:
    """Builds under an exclusive file lock (test processes running in
    parallel each ask for the build) and renames the library into place, so
    no process loads a half-written one."""
    out_dir = HERE / "_build"
    out_dir.mkdir(exist_ok=True)
    with open(out_dir / ".lock", "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        return _build(out_dir, asan)


def _build(out_dir: Path, asan: bool) -> Path:
    flags = ["-std=c++20", "-O1", "-g", "-fPIC", "-pthread", "-ffp-contract=off",
             f"-I{HERE / 'include'}", f"-I{ROOT / 'include'}", "-Wno-attributes",
             "-DPS_ENABLE_BATCH"]
    if asan:
        flags += ["-fsanitize=address", "-fno-omit-frame-pointer"]
    objs = []

    def cc(src: Path, lang_cxx: bool) -> Path:
        obj = out_dir / (src.stem + (".asan" if asan else "") + ".o")
        cmd = ["g++", *flags, *(["-x", "c++"] if lang_cxx else []), "-c", str(src), "-o", str(obj)]
        subprocess.run(cmd, check=True)
        return obj

    jobs = [(CSRC / s, True) for s in SOURCES] + [(HERE / "emu_runtime.cpp", False)]
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda j: cc(*j), jobs))
    lib = out_dir / ("libpolysplat_emu_asan.so" if asan else "libpolysplat_emu.so")
    tmp = lib.with_name(lib.name + f".tmp{os.getpid()}")
    subprocess.run(["g++", "-shared", "-pthread", *(["-fsanitize=address"] if asan else []),
                    *map(str, objs), "-o", str(tmp)], check=True)
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--asan", action="store_true")
    print(build(ap.parse_args().asan))
    sys.exit(0)
