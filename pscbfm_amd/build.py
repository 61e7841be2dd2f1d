"""Build recipe of libpscbfm.so (HIP kernels for gfx950 + C++ host engine).

    python -m pscbfm_amd.build          # or pscbfm_amd.build.build()

Compiles every source in pscbfm_amd/csrc with hipcc into objects under
pscbfm_amd/_build and links pscbfm_amd/libpscbfm.so in-tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
OUT = HERE / "libpscbfm.so"
BUILD = HERE / "_build"
ARCH = os.environ.get("PSCBFM_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]
# Device code: gfx950 only; packed-FP32 instruction selection is disabled.
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HOST = ["-x", "c++", "-D__HIP_PLATFORM_AMD__=1", "-isystem", f"{ROCM}/include"]


def _hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found")
    return path


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    flags = DEVICE if src.suffix == ".hip" else HOST
    cmd = [_hipcc(), *COMMON, *flags, "-I", str(CSRC), "-I", str(INCLUDE), "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if res.stderr.strip():
        sys.stderr.write(_filter_warnings(res.stderr))
    return obj


def _filter_warnings(text: str) -> str:
    # hipcc's host pass does not know the packed-fp32 feature; that notice is harmless.
    keep = [ln for ln in text.splitlines(keepends=True) if "packed-fp32-ops" not in ln]
    return "".join(keep)


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted([*CSRC.glob("*.hip"), *CSRC.glob("*.cpp")])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [_hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", *map(str, objs), "-lz", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, OUT)
    # command-line simulator over the C ABI (host code only)
    cli_src = HERE / "cli" / "pscbfm_run.cpp"
    cli = HERE / "pscbfm_run"
    cmd = [shutil.which("g++") or "g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", str(INCLUDE), str(cli_src),
           "-o", str(cli), "-L", str(HERE), "-lpscbfm", f"-Wl,-rpath,$ORIGIN", f"-Wl,-rpath-link,{ROCM}/lib",
           "-Wl,--allow-shlib-undefined"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"CLI build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {OUT} and {cli}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
