"""ctypes front-end of the FP64 CPU oracle (ppf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.

Array conventions (numpy, C order, x fastest in memory as in the reference's
Fortran arrays):
  rhs       shape (n3m, n2m, n1m)
  P         shape (n3m+2, n2m+2, n1m+2)      one ghost layer per side
  x, dx, dmx length nm+2                     Fortran index 0..nm+1
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libppf_oracle.so")
_REF_TDMAS = os.path.join(_HERE, "_ref", "libref_tdmas.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")


def build(force: bool = False) -> None:
    """Compile the oracle (and the reference Thomas where the reference exists)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "oracle"], check=True)


def _load() -> C.CDLL:
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    I, D = C.c_int, C.c_double
    lib.or_grid.argtypes = [I, I, D, D, I, _dp, _dp, _dp]
    lib.or_eigen.argtypes = [I, I, I, I, D, D, _dp, _dp]
    lib.or_fft.argtypes = [I, _cp, I]
    lib.or_rfft.argtypes = [I, _dp, _cp]
    lib.or_irfft.argtypes = [I, _cp, _dp]
    lib.or_dct2.argtypes = [I, _dp, _dp]
    lib.or_idct2.argtypes = [I, _dp, _dp]
    lib.or_tdma_many.argtypes = [_dp, _dp, _dp, _dp, C.c_long, I]
    lib.or_tdma_cycl_many.argtypes = [_dp, _dp, _dp, _dp, C.c_long, I]
    lib.or_solve.argtypes = [I, I, I, D, D, D, I, I, I, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.or_solve.restype = I
    lib.or_apply_L.argtypes = [I, I, I, D, D, I, I, _dp, _dp, _dp, _dp]
    lib.or_rel_residual.argtypes = [I, I, I, D, D, I, I, _dp, _dp, _dp, _dp]
    lib.or_rel_residual.restype = D
    lib.or_num_threads.restype = I
    lib.or_set_threads.argtypes = [I]
    return lib


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def num_threads() -> int:
    return int(lib().or_num_threads())


def grid(uniform: bool, nm: int, L: float, gamma: float, pbc: bool):
    x = np.zeros(nm + 2)
    dx = np.zeros(nm + 2)
    dmx = np.zeros(nm + 2)
    lib().or_grid(int(bool(uniform)), nm, float(L), float(gamma), int(bool(pbc)), x, dx, dmx)
    return x, dx, dmx


def eigen(pbc1: bool, pbc2: bool, n1m: int, n2m: int, L1: float, L2: float):
    dxk2 = np.zeros(n1m)
    dyk2 = np.zeros(n2m)
    lib().or_eigen(int(pbc1), int(pbc2), n1m, n2m, float(L1), float(L2), dxk2, dyk2)
    return dxk2, dyk2


def fft(z, sign: int = -1):
    z = np.ascontiguousarray(z, dtype=np.complex128).copy()
    lib().or_fft(z.size, z, int(sign))
    return z


def rfft(x):
    x = _f64(x)
    out = np.zeros(x.size // 2 + 1, dtype=np.complex128)
    lib().or_rfft(x.size, x, out)
    return out


def irfft(X, n: int):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    out = np.zeros(n)
    lib().or_irfft(n, X, out)
    return out


def dct2(x):
    x = _f64(x)
    out = np.zeros_like(x)
    lib().or_dct2(x.size, x, out)
    return out


def idct2(X):
    X = _f64(X)
    out = np.zeros_like(X)
    lib().or_idct2(X.size, X, out)
    return out


def tdma_many(a, b, c, d):
    """Thomas on arrays of shape (nrow, nsys); returns the solution (inputs untouched)."""
    a, b = _f64(a), _f64(b)
    c, d = _f64(c).copy(), _f64(d).copy()
    nrow, nsys = d.shape
    lib().or_tdma_many(a, b, c, d, nsys, nrow)
    return d


def tdma_cycl_many(a, b, c, d):
    """Cyclic Thomas on arrays of shape (nrow, nsys); returns the solution."""
    a, b = _f64(a), _f64(b)
    c, d = _f64(c).copy(), _f64(d).copy()
    nrow, nsys = d.shape
    lib().or_tdma_cycl_many(a, b, c, d, nsys, nrow)
    return d


def ref_tdma_available() -> bool:
    return os.path.exists(_REF_TDMAS)


def ref_tdma_cycl_many(a, b, c, d):
    """The reference's own tdma_cycl_many (00_C/PaScaL_TDMA/src/tdmas.cpp:161-210),
    compiled from source into oracle/_ref.  It strides rows by n2 (the row
    count) where tdma_many strides by n1, so it agrees with the Fortran
    tdma_cycl_many layout only for square batches (nsys == nrow)."""
    rl = C.CDLL(_REF_TDMAS)
    rl.tdma_cycl_many.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int]
    a, b = _f64(a).copy(), _f64(b).copy()
    c, d = _f64(c).copy(), _f64(d).copy()
    nrow, nsys = d.shape
    if nrow != nsys:
        raise ValueError("the reference tdma_cycl_many is layout-consistent only for nsys == nrow")
    rl.tdma_cycl_many(a, b, c, d, nsys, nrow)
    return d


def ref_tdma_many(a, b, c, d):
    """The reference's own tdma_many (00_C/PaScaL_TDMA/src/tdmas.cpp:79-103),
    compiled from source into oracle/_ref; same array convention as tdma_many."""
    rl = C.CDLL(_REF_TDMAS)
    rl.tdma_many.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int]
    a, b = _f64(a).copy(), _f64(b).copy()
    c, d = _f64(c).copy(), _f64(d).copy()
    nrow, nsys = d.shape
    rl.tdma_many(a, b, c, d, nsys, nrow)
    return d


def solve(rhs, L, pbc, dmx1, dmx2, dmx3, dx3):
    """The reference solve (cuda_Poisson_FFT_1D) restated in FP64 on the CPU."""
    rhs = _f64(rhs)
    n3m, n2m, n1m = rhs.shape
    P = np.zeros((n3m + 2, n2m + 2, n1m + 2))
    rc = lib().or_solve(n1m, n2m, n3m, float(L[0]), float(L[1]), float(L[2]),
                        int(pbc[0]), int(pbc[1]), int(pbc[2]), rhs, P,
                        _f64(dmx1), _f64(dmx2), _f64(dmx3), _f64(dx3))
    if rc != 0:
        raise ValueError(f"oracle solve rejected the problem (code {rc})")
    return P


def apply_L(p, L, pbc, dmx3, dx3):
    p = _f64(p)
    n3m, n2m, n1m = p.shape
    out = np.zeros_like(p)
    lib().or_apply_L(n1m, n2m, n3m, float(L[0]), float(L[1]), int(pbc[0]), int(pbc[1]),
                     p, out, _f64(dmx3), _f64(dx3))
    return out


def rel_residual(f, P, L, pbc, dmx3, dx3):
    f = _f64(f)
    n3m, n2m, n1m = f.shape
    return float(lib().or_rel_residual(n1m, n2m, n3m, float(L[0]), float(L[1]), int(pbc[0]), int(pbc[1]),
                                       f, _f64(P), _f64(dmx3), _f64(dx3)))
