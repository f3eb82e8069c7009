"""Builds the package's shared libraries in-tree: the engine with hipcc for
gfx950, the trajectory frame reader (host code) with the C++ compiler.

``build_library()`` and ``build_xdr_library()`` always recompile; the
loaders (``pycusaxs_amd._lib``, ``pycusaxs_amd.trajectory``) only load what
is built and never compile on their own.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SRC_DIR = PKG / "csrc"
LIB_DIR = PKG / "_lib"
LIB_PATH = LIB_DIR / "libsaxs_engine.so"
SOURCES = [SRC_DIR / "saxs_engine.hip"]
XDR_LIB_PATH = LIB_DIR / "libsaxs_xdr.so"
XDR_SOURCES = [SRC_DIR / "xdr_frames.cpp"]
ARCH = os.environ.get("PYCUSAXS_AMD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the MI355X engine cannot be built")


def build_command(out: Path = LIB_PATH):
    return [
        hipcc(), "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
        # packed fp32 instructions are kept off: scalar fp32 code only
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics",
        f"-I{ROOT / 'include'}", f"-I{SRC_DIR}",
        "-o", str(out), *[str(s) for s in SOURCES],
    ]


def build_library(verbose: bool = False) -> Path:
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = build_command(tmp)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose and res.stderr:
        print(res.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def cxx() -> str:
    for cand in (os.environ.get("CXX"), shutil.which("g++"), shutil.which("c++")):
        if cand and (Path(cand).exists() or shutil.which(cand)):
            return cand
    raise RuntimeError("no C++ compiler found for the trajectory reader")


def build_xdr_library() -> Path:
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    tmp = XDR_LIB_PATH.with_suffix(".so.tmp")
    cmd = [cxx(), "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread",
           f"-I{ROOT / 'include'}", "-o", str(tmp), *[str(s) for s in XDR_SOURCES]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("C++ build of the trajectory reader failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, XDR_LIB_PATH)
    return XDR_LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_xdr_library())
