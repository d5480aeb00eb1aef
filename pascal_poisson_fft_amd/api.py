"""Python host side of the solve path, over the C ABI (include/pascal_poisson_fft.h).

Mirrors the reference's module interface (01_Fortran/src/fft_3d_poisson.f90):

    reference                                   here
    fft_poisson_plan_cuda_create(...) :70       FFTPoisson(n1m, n2m, n3m, L, pbc, ...)
    cuda_Poisson_FFT_initial()        :110      (done by the constructor)
    cuda_Poisson_FFT_1D(PRHS_d, P_d, dmx1_d,
        dmx2_d, dmx3_d, dx3_d)        :353      FFTPoisson.solve(prhs, p, dmx1, dmx2, dmx3, dx3)
    cuda_Poisson_FFT_clean()          :123      FFTPoisson.clean()

Device memory comes from the library's own shim (DeviceArray), so the process
carries exactly one HIP runtime.  The library must be built
(python -m pascal_poisson_fft_amd.build); loading fails loudly otherwise --
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libpascal_poisson_fft.so")

PPF_PREC_FP64 = 0
PPF_PREC_FP32_IR = 1
PPF_ERR_INVALID_ARG = -1
PPF_ERR_SIZE = -2
PPF_ERR_HIP = -3
PPF_ERR_ALLOC = -4
PPF_ERR_UNSUPPORTED = -5
PPF_ERR_NOT_CONVERGED = -6
PPF_ERR_SCRATCH = -7

H2D, D2H, D2D = 1, 2, 3


class PPFError(RuntimeError):
    def __init__(self, code: int, what: str):
        self.code = code
        super().__init__(f"{what}: {error_string(code)} ({code})")


class Desc(C.Structure):
    _fields_ = [
        ("n1m", C.c_int), ("n2m", C.c_int), ("n3m", C.c_int),
        ("L1", C.c_double), ("L2", C.c_double), ("L3", C.c_double),
        ("pbc1", C.c_int), ("pbc2", C.c_int), ("pbc3", C.c_int),
        ("precision", C.c_int), ("tol", C.c_double), ("max_sweeps", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [("sweeps", C.c_int), ("rel_residual", C.c_double)]


class BatchIO(C.Structure):
    _fields_ = [("prhs", C.c_void_p), ("p", C.c_void_p), ("dmx1", C.c_void_p), ("dmx2", C.c_void_p),
                ("dmx3", C.c_void_p), ("dx3", C.c_void_p)]


PPF_BATCH_NO_GRAPH = 1
PPF_BATCH_NO_TUNE = 2


_lib = None


def lib() -> C.CDLL:
    """Load the gfx950 library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m pascal_poisson_fft_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, sz, I, D = C.c_void_p, C.c_size_t, C.c_int, C.c_double
    L.ppf_plan_create.argtypes = [C.POINTER(Desc), C.POINTER(vp)]
    L.ppf_solve.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(Stats)]
    L.ppf_solve_timed.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(C.c_float)]
    L.ppf_plan_destroy.argtypes = [vp]
    L.ppf_residual.argtypes = [vp, vp, vp, vp, vp, vp, C.POINTER(D)]
    L.ppf_plan_workspace_bytes.argtypes = [vp]
    L.ppf_plan_workspace_bytes.restype = sz
    L.ppf_error_string.argtypes = [I]
    L.ppf_error_string.restype = C.c_char_p
    L.ppf_tdma_plan_many_create.argtypes = [I, I, I, C.POINTER(vp)]
    L.ppf_tdma_many_solve.argtypes = [vp, vp, vp, vp, vp, vp]
    L.ppf_tdma_plan_many_destroy.argtypes = [vp]
    L.ppf_tdma_many_solve_cycle.argtypes = [vp, vp, vp, vp, vp, vp]
    L.ppf_malloc.argtypes = [C.POINTER(vp), sz]
    L.ppf_free.argtypes = [vp]
    L.ppf_memcpy.argtypes = [vp, vp, sz, I]
    L.ppf_memset.argtypes = [vp, I, sz]
    L.ppf_sync.argtypes = [vp]
    L.ppf_device_count.argtypes = [C.POINTER(I)]
    L.ppf_set_device.argtypes = [I]
    L.ppf_copy_timed.argtypes = [vp, vp, sz, I, vp, C.POINTER(C.c_float)]
    L.ppf_solve_sweeps.argtypes = [vp, vp, vp, vp, vp, vp, vp, I, vp]
    L.ppf_solve_async.argtypes = [vp, vp, vp, vp, vp, vp, vp, I, I, vp]
    L.ppf_solve_noalloc.argtypes = [C.POINTER(Desc), vp, vp, vp, vp, vp, vp, vp]
    L.ppf_solve_async_result.argtypes = [vp, vp, C.POINTER(Stats)]
    L.ppf_plan_tune.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(I)]
    L.ppf_plan_set_zvariant.argtypes = [vp, I]
    L.ppf_plan_variants.argtypes = [vp, C.POINTER(I)]
    L.ppf_plan_set_variants.argtypes = [vp, I, I, I]
    L.ppf_batch_create.argtypes = [I, C.POINTER(Desc), C.POINTER(BatchIO), I, C.POINTER(vp)]
    L.ppf_batch_solve.argtypes = [vp, vp, C.POINTER(Stats)]
    L.ppf_batch_destroy.argtypes = [vp]
    L.ppf_synth_divergence_rhs.argtypes = [vp, I, I, I, D, D, vp, vp, I, I, C.c_ulonglong, D, vp]
    _lib = L
    return L


def error_string(code: int) -> str:
    return lib().ppf_error_string(int(code)).decode()


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise PPFError(rc, what)


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().ppf_device_count(C.byref(n)), "ppf_device_count")
    return n.value


def sync(stream=None) -> None:
    _check(lib().ppf_sync(stream), "ppf_sync")


def synth_divergence_rhs(f: "DeviceArray", n1m: int, n2m: int, n3m: int, L1: float, L2: float,
                         dx3: "DeviceArray", tables: "DeviceArray", nmodes: int, stride: int, seed: int,
                         noise: float, stream=None) -> None:
    """Projection-step RHS: divergence of a synthetic staggered velocity field
    (ppf_synth_divergence_rhs)."""
    _check(lib().ppf_synth_divergence_rhs(f.ptr, n1m, n2m, n3m, float(L1), float(L2), dx3.ptr,
                                          tables.ptr if tables is not None else None, int(nmodes), int(stride),
                                          int(seed), float(noise), stream), "ppf_synth_divergence_rhs")


def copy_bandwidth(count: int = 1 << 28, reps: int = 10) -> float:
    """GB/s of the library's plain device copy kernel over `count` doubles
    (read + write bytes / time)."""
    a, b = DeviceArray(count), DeviceArray(count)
    try:
        a.fill_zero()
        ms = C.c_float(0.0)
        _check(lib().ppf_copy_timed(b.ptr, a.ptr, count, 1, None, C.byref(ms)), "ppf_copy_timed")
        _check(lib().ppf_copy_timed(b.ptr, a.ptr, count, reps, None, C.byref(ms)), "ppf_copy_timed")
        return 16.0 * count / (ms.value * 1e-3) / 1e9
    finally:
        a.free()
        b.free()


def solve_noalloc(n1m: int, n2m: int, n3m: int, L, pbc, prhs: "DeviceArray", p: "DeviceArray",
                  dmx1: "DeviceArray", dmx2: "DeviceArray", dmx3: "DeviceArray", dx3: "DeviceArray",
                  stream=None) -> None:
    """FP64 solve without a plan and without device allocation (ppf_solve_noalloc):
    the library's static device scratch holds the workspace; calls on one GPU
    must not overlap."""
    d = Desc(int(n1m), int(n2m), int(n3m), float(L[0]), float(L[1]), float(L[2]),
             int(bool(pbc[0])), int(bool(pbc[1])), int(bool(pbc[2])), PPF_PREC_FP64, 0.0, 0)
    _check(lib().ppf_solve_noalloc(C.byref(d), prhs.ptr, p.ptr, dmx1.ptr, dmx2.ptr, dmx3.ptr, dx3.ptr, stream),
           "ppf_solve_noalloc")


class DeviceArray:
    """A device buffer from ppf_malloc with a numpy shape/dtype for transfers."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(lib().ppf_malloc(C.byref(p), self.nbytes), "ppf_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _check(lib().ppf_memcpy(self.ptr, a.ctypes.data, self.nbytes, H2D), "ppf_memcpy H2D")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _check(lib().ppf_memcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "ppf_memcpy D2H")
        return out

    def fill_zero(self) -> None:
        _check(lib().ppf_memset(self.ptr, 0, self.nbytes), "ppf_memset")

    def free(self) -> None:
        if self.ptr:
            lib().ppf_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class FFTPoisson:
    """One plan of the solve path (module fft_poisson of the reference)."""

    def __init__(self, n1m: int, n2m: int, n3m: int, L=(2.0, 2.0, 2.0), pbc=(False, False, False),
                 precision: int = PPF_PREC_FP64, tol: float = 1e-10, max_sweeps: int = 8):
        self.desc = Desc(int(n1m), int(n2m), int(n3m), float(L[0]), float(L[1]), float(L[2]),
                         int(bool(pbc[0])), int(bool(pbc[1])), int(bool(pbc[2])),
                         int(precision), float(tol), int(max_sweeps))
        h = C.c_void_p()
        _check(lib().ppf_plan_create(C.byref(self.desc), C.byref(h)), "ppf_plan_create")
        self.handle = h.value

    @property
    def shape(self):
        return (self.desc.n1m, self.desc.n2m, self.desc.n3m)

    def workspace_bytes(self) -> int:
        return int(lib().ppf_plan_workspace_bytes(self.handle))

    def solve(self, prhs: DeviceArray, p: DeviceArray, dmx1: DeviceArray, dmx2: DeviceArray, dmx3: DeviceArray,
              dx3: DeviceArray, stream=None, allow_not_converged: bool = False) -> Stats:
        """One solve; returns its Stats (FP32_IR: sweeps and final relative
        residual).  A refinement that stops above tol raises, unless
        allow_not_converged (p then holds the last iterate)."""
        st = Stats()
        rc = lib().ppf_solve(self.handle, prhs.ptr, p.ptr, dmx1.ptr, dmx2.ptr, dmx3.ptr, dx3.ptr, stream, C.byref(st))
        if not (rc == PPF_ERR_NOT_CONVERGED and allow_not_converged):
            _check(rc, "ppf_solve")
        return st

    STAGES = ("x-forward", "y-forward", "z-thomas", "y-inverse", "x-inverse")

    def solve_timed(self, prhs: DeviceArray, p: DeviceArray, dmx1: DeviceArray, dmx2: DeviceArray,
                    dmx3: DeviceArray, dx3: DeviceArray, stream=None) -> list:
        """One solve with HIP events around each pass; returns the 5 pass times in ms."""
        ms = (C.c_float * 5)()
        _check(lib().ppf_solve_timed(self.handle, prhs.ptr, p.ptr, dmx1.ptr, dmx2.ptr, dmx3.ptr, dx3.ptr, stream,
                                     ms), "ppf_solve_timed")
        return [float(v) for v in ms]

    def residual(self, f: DeviceArray, p: DeviceArray, dmx3: DeviceArray, dx3: DeviceArray, stream=None) -> float:
        r = C.c_double()
        _check(lib().ppf_residual(self.handle, f.ptr, p.ptr, dmx3.ptr, dx3.ptr, stream, C.byref(r)), "ppf_residual")
        return r.value

    def solve_sweeps(self, prhs: DeviceArray, p: DeviceArray, dmx1: DeviceArray, dmx2: DeviceArray,
                     dmx3: DeviceArray, dx3: DeviceArray, sweeps: int, stream=None) -> None:
        """Exactly `sweeps` refinement corrections, no convergence test (capture-safe)."""
        _check(lib().ppf_solve_sweeps(self.handle, prhs.ptr, p.ptr, dmx1.ptr, dmx2.ptr, dmx3.ptr, dx3.ptr,
                                      int(sweeps), stream), "ppf_solve_sweeps")

    def solve_async(self, prhs: DeviceArray, p: DeviceArray, dmx1: DeviceArray, dmx2: DeviceArray,
                    dmx3: DeviceArray, dx3: DeviceArray, min_sweeps: int = 0, extra_sweeps: int = -1,
                    stream=None) -> None:
        """Solve with the FP32_IR convergence test on the device (no host synchronisation)."""
        _check(lib().ppf_solve_async(self.handle, prhs.ptr, p.ptr, dmx1.ptr, dmx2.ptr, dmx3.ptr, dx3.ptr,
                                     int(min_sweeps), int(extra_sweeps), stream), "ppf_solve_async")

    def async_result(self, stream=None, allow_not_converged: bool = False) -> Stats:
        """Sweeps and residual of the last solve_async (synchronises)."""
        st = Stats()
        rc = lib().ppf_solve_async_result(self.handle, stream, C.byref(st))
        if not (rc == PPF_ERR_NOT_CONVERGED and allow_not_converged):
            _check(rc, "ppf_solve_async_result")
        return st

    def tune(self, prhs: DeviceArray, p: DeviceArray, dmx1: DeviceArray, dmx2: DeviceArray, dmx3: DeviceArray,
             dx3: DeviceArray, stream=None) -> int:
        """Autotune the z-pass shape on this data; returns the chosen id."""
        v = C.c_int(0)
        _check(lib().ppf_plan_tune(self.handle, prhs.ptr, p.ptr, dmx1.ptr, dmx2.ptr, dmx3.ptr, dx3.ptr, stream,
                                   C.byref(v)), "ppf_plan_tune")
        return v.value

    def set_zvariant(self, v: int) -> None:
        _check(lib().ppf_plan_set_zvariant(self.handle, int(v)), "ppf_plan_set_zvariant")

    def variants(self) -> tuple:
        """(x-pass variant, y-pass variant, z shape) the plan runs."""
        v = (C.c_int * 3)()
        _check(lib().ppf_plan_variants(self.handle, v), "ppf_plan_variants")
        return tuple(int(x) for x in v)

    def set_variants(self, xv: int, yv: int, zv: int) -> None:
        _check(lib().ppf_plan_set_variants(self.handle, int(xv), int(yv), int(zv)), "ppf_plan_set_variants")

    def clean(self) -> None:
        if self.handle:
            _check(lib().ppf_plan_destroy(self.handle), "ppf_plan_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.clean()
        except Exception:
            pass


class Batch:
    """Many independent problems bound to their device buffers (ppf_batch_*)."""

    def __init__(self, problems, flags: int = 0):
        """problems: list of (FFTPoisson-style desc kwargs dict, Problem3D)."""
        n = len(problems)
        descs = (Desc * n)()
        ios = (BatchIO * n)()
        for i, (kw, dev) in enumerate(problems):
            L = kw.get("L", (2.0, 2.0, 2.0))
            pbc = kw.get("pbc", (False, False, False))
            descs[i] = Desc(int(kw["n1m"]), int(kw["n2m"]), int(kw["n3m"]), float(L[0]), float(L[1]), float(L[2]),
                            int(bool(pbc[0])), int(bool(pbc[1])), int(bool(pbc[2])),
                            int(kw.get("precision", PPF_PREC_FP64)), float(kw.get("tol", 1e-10)),
                            int(kw.get("max_sweeps", 8)))
            ios[i] = BatchIO(dev.rhs.ptr, dev.p.ptr, dev.dmx1.ptr, dev.dmx2.ptr, dev.dmx3.ptr, dev.dx3.ptr)
        h = C.c_void_p()
        _check(lib().ppf_batch_create(n, descs, ios, int(flags), C.byref(h)), "ppf_batch_create")
        self.handle = h.value
        self.n = n

    def solve(self, stream=None, stats: bool = False):
        """Solve every problem; with stats, returns [(sweeps, rel_residual)] (synchronises)."""
        st = (Stats * self.n)() if stats else None
        _check(lib().ppf_batch_solve(self.handle, stream, st), "ppf_batch_solve")
        return [(s.sweeps, s.rel_residual) for s in st] if stats else None

    def destroy(self) -> None:
        if self.handle:
            lib().ppf_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class TDMAPlanMany:
    """PaScaL_TDMA_plan_many_create_cuda / PaScaL_TDMA_many_solve_cuda (one rank)."""

    def __init__(self, nx_sys: int, ny_sys: int, nz_row: int):
        h = C.c_void_p()
        _check(lib().ppf_tdma_plan_many_create(int(nx_sys), int(ny_sys), int(nz_row), C.byref(h)),
               "ppf_tdma_plan_many_create")
        self.handle = h.value

    def solve(self, a: DeviceArray, b: DeviceArray, c: DeviceArray, d: DeviceArray, stream=None) -> None:
        _check(lib().ppf_tdma_many_solve(self.handle, a.ptr, b.ptr, c.ptr, d.ptr, stream), "ppf_tdma_many_solve")

    def solve_cycle(self, a: DeviceArray, b: DeviceArray, c: DeviceArray, d: DeviceArray, stream=None) -> None:
        """PaScaL_TDMA_many_solve_cycle_cuda (one rank): cyclic systems."""
        _check(lib().ppf_tdma_many_solve_cycle(self.handle, a.ptr, b.ptr, c.ptr, d.ptr, stream),
               "ppf_tdma_many_solve_cycle")

    def destroy(self) -> None:
        if self.handle:
            lib().ppf_tdma_plan_many_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Problem3D:
    """Device-resident inputs of one problem: RHS, solution buffer and metrics."""

    def __init__(self, rhs: np.ndarray, ax, ay, az):
        n3m, n2m, n1m = rhs.shape
        self.rhs = DeviceArray.from_host(rhs)
        self.p = DeviceArray((n3m + 2, n2m + 2, n1m + 2))
        self.p.fill_zero()
        self.dmx1 = DeviceArray.from_host(ax.dmx)
        self.dmx2 = DeviceArray.from_host(ay.dmx)
        self.dmx3 = DeviceArray.from_host(az.dmx)
        self.dx3 = DeviceArray.from_host(az.dx)

    def free(self):
        for a in (self.rhs, self.p, self.dmx1, self.dmx2, self.dmx3, self.dx3):
            a.free()


def solve_host(rhs: np.ndarray, ax, ay, az, precision: int = PPF_PREC_FP64, **kw) -> np.ndarray:
    """Convenience: upload, solve once on the GPU, download P (n3m+2, n2m+2, n1m+2)."""
    n3m, n2m, n1m = rhs.shape
    plan = FFTPoisson(n1m, n2m, n3m, (ax.L, ay.L, az.L), (ax.pbc, ay.pbc, az.pbc), precision=precision, **kw)
    prob = Problem3D(rhs, ax, ay, az)
    plan.solve(prob.rhs, prob.p, prob.dmx1, prob.dmx2, prob.dmx3, prob.dx3)
    sync()
    out = prob.p.download()
    prob.free()
    plan.clean()
    return out
