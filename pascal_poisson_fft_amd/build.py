"""Build the gfx950 library (and the Fortran interface) in-tree.

    python -m pascal_poisson_fft_amd.build            # library + Fortran module/driver
    python -m pascal_poisson_fft_amd.build --lib-only

Outputs (git-ignored, shipped with the source tree):
    pascal_poisson_fft_amd/lib/libpascal_poisson_fft.so
    pascal_poisson_fft_amd/lib/fft_poisson.mod, pascal_poisson_fft_amd/lib/ppf_example
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "_build")
LIBNAME = "libpascal_poisson_fft.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("PPF_FC", "/opt/rocm/bin/amdflang")
ARCH = os.environ.get("PPF_OFFLOAD_ARCH", "gfx950")

HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall",
    # packed FP32 instructions stay off: the compiler is not allowed to emit them
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *HIPFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build_lib(jobs: int = 8) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(_compile, srcs))
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)
    return out


def build_fortran() -> None:
    """Compile the Fortran module fft_poisson and the example driver with amdflang."""
    fdir = os.path.join(PKG, "fortran")
    if not os.path.isdir(fdir) or not os.path.exists(FC):
        return
    srcs = [os.path.join(fdir, f) for f in ("fft_poisson.f90", "example_main.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return
    os.makedirs(LIBDIR, exist_ok=True)
    exe = os.path.join(LIBDIR, "ppf_example")
    cmd = [FC, "-O2", "-J", LIBDIR, *srcs, "-L", LIBDIR, "-lpascal_poisson_fft",
           "-Wl,-rpath,$ORIGIN", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=LIBDIR)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{r.stderr}")


XCHECK_SRC = os.path.join(PKG, "xcheck", "rocfft_xcheck.hip")
XCHECK_LIB = "libppf_rocfft_xcheck.so"


def build_xcheck() -> str | None:
    """The vendor-FFT timing cross-check (bench only; the solve library does
    not link it).  Best effort: a missing hipFFT skips it."""
    if not os.path.exists(XCHECK_SRC):
        return None
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, XCHECK_LIB)
    cmd = [HIPCC, *HIPFLAGS, "-shared", XCHECK_SRC, "-o", out, "-lhipfft"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(f"rocFFT cross-check not built:\n{r.stderr[-2000:]}", file=sys.stderr)
        return None
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    out = build_lib(a.j)
    print("built", out)
    if not a.lib_only:
        build_fortran()
        build_xcheck()


if __name__ == "__main__":
    sys.exit(main())
