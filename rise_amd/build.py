"""Builds the HIP shared library ``rise_amd/librise_hip.so`` for gfx950.

Every translation unit under ``rise_amd/csrc`` is compiled with hipcc, in
parallel, and linked into one library that exports the C ABI declared in
``include/*.h``.  Flags:

* ``--offload-arch=gfx950``: MI355X only.
* ``-ffp-contract=off``: double-precision intersection arithmetic is rounded
  operation by operation, exactly as the reference's C++ writes it.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 ops.

Run ``python -m rise_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"
LIB = PKG / "librise_hip.so"
DEMO = PKG / "adapter_demo"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
    "-Wall",
    "-Wno-unused-result",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    cmd = [HIPCC, *COMMON]
    if src.suffix == ".hip":
        cmd += ["-x", "hip", *DEVICE]
    else:
        # Host-only translation units: plain C++ with the HIP runtime headers.
        cmd += ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
    cmd += ["-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compiles every source and links ``librise_hip.so``; returns its path."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(parents=True, exist_ok=True)
    srcs = _sources()
    jobs = min(len(srcs), max(1, min(16, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    _build_demo()
    if verbose:
        print(f"built {LIB} and {DEMO}")
    return LIB


def _build_demo() -> None:
    """The C++ adapter demo (rise_amd/host), linked against the library."""
    src = PKG / "host" / "adapter_demo.cpp"
    cmd = [HIPCC, "-x", "c++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{INCLUDE}", f"-I{PKG / 'host'}", str(src), "-o", str(DEMO), f"-L{PKG}", "-lrise_hip",
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"demo build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
