"""Build ``libvjx.so`` (HIP kernels for gfx950 + host C++) in-tree.

Usage: ``python -m vajax_amd.build`` or ``vajax_amd.build.build()``.  Objects
go to ``_build/`` at the repository root; the shared library is written next
to this file so it travels with the source tree.

``python -m vajax_amd.build --asan-host DIR`` (``--ubsan-host DIR``) builds, into
DIR, a variant whose host C++ (analysis, orderings, C ABI, drivers) is
compiled with AddressSanitizer (UndefinedBehaviorSanitizer), the device code
unchanged (build the normal library first: its kernel objects are reused).
The CPU tests run against it with ``VJX_LIB=DIR/libvjx.so`` and the
sanitizer runtime preloaded (``LD_PRELOAD=<clang resource
dir>/lib/linux/libclang_rt.asan-x86_64.so``, or ``ubsan_standalone``).
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
ROOT = PKG.parent
OBJ = ROOT / "_build"
LIB = PKG / "libvjx.so"

ARCH = os.environ.get("VJX_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: FP64 expressions round as written (matches the numpy
# oracle); packed FP32 ops are disabled for correctness on this target.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
CXX_FLAGS = ["-x", "c++", "-O3", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__",
             "-I/opt/rocm/include", "-ffp-contract=off"]

HIP_SOURCES = ["kernels.hip", "transient_kernels.hip", "noalloc.hip", "dense.hip"]
CXX_SOURCES = ["symbolic.cpp", "separator.cpp", "api.cpp", "lu.cpp", "circuit.cpp", "transient.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        if (CSRC / src).exists():
            jobs.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(OBJ / (src + ".o"))])
    for src in CXX_SOURCES:
        jobs.append([HIPCC, *CXX_FLAGS, "-c", str(CSRC / src), "-o", str(OBJ / (src + ".o"))])
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr, file=sys.stderr)
    objs = [str(OBJ / (s + ".o")) for s in HIP_SOURCES + CXX_SOURCES if (CSRC / s).exists()]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


def build_asan_host(out_dir: str, sanitizer: str = "address") -> Path:
    """The host sources under a sanitizer ("address" or "undefined";
    -Xarch_host: no device code is instrumented), linked with the kernel
    objects of the last normal build into ``out_dir/libvjx.so``."""
    if sanitizer not in ("address", "undefined"):
        raise ValueError("sanitizer must be 'address' or 'undefined'")
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    flags = [f for f in CXX_FLAGS if f != "-O3"] + [
        "-O1", "-g", "-fno-omit-frame-pointer", "-Xarch_host", f"-fsanitize={sanitizer}"]
    objs = [str(OBJ / (s + ".o")) for s in HIP_SOURCES if (CSRC / s).exists()]
    missing = [o for o in objs if not Path(o).exists()]
    if missing:
        raise RuntimeError(f"build the library first (missing {missing[0]})")
    jobs = [[HIPCC, *flags, "-c", str(CSRC / src), "-o", str(out / (src + ".o"))]
            for src in CXX_SOURCES]
    with ThreadPoolExecutor(max_workers=min(len(jobs), 8)) as ex:
        list(ex.map(_run, jobs))
    objs += [str(out / (s + ".o")) for s in CXX_SOURCES]
    lib = out / "libvjx.so"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-Xarch_host",
          f"-fsanitize={sanitizer}", "-shared-libsan", *objs, "-o", str(lib)])
    return lib


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] in ("--asan-host", "--ubsan-host"):
        kind = "address" if sys.argv[1] == "--asan-host" else "undefined"
        print(build_asan_host(sys.argv[2], kind))
    else:
        print(build(verbose=True))
