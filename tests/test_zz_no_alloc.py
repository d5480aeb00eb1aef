"""The allocation-free solve (ppf_solve_noalloc) against the plan-based solve
(ppf_solve), through the C ABI.

Two calls in a row share the library's static device scratch: a larger
problem first, then a smaller one of other boundary conditions, so stale
contents of the first call would show in the second.  Each must match the
plan-based solve of the same problem to <= 1e-12 relative L2 of the ghosted
solution (the FP64 parity bar); the axis tables are built on the device here
and on the host for plans, so the two agree to rounding, not bitwise.  A
problem beyond the 64 MiB scratch is rejected with PPF_ERR_SCRATCH, and an
FP32_IR description with PPF_ERR_UNSUPPORTED.
"""
import ctypes

import numpy as np
import pytest

from pascal_poisson_fft_amd import api
from pascal_poisson_fft_amd import workloads as W

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _both(prob):
    ax, ay, az = prob.ax
    n1, n2, n3 = prob.shape
    dev = api.Problem3D(prob.rhs, ax, ay, az)
    try:
        plan = api.FFTPoisson(n1, n2, n3, prob.L, prob.pbc)
        plan.solve(dev.rhs, dev.p, dev.dmx1, dev.dmx2, dev.dmx3, dev.dx3)
        api.sync()
        want = dev.p.download()
        plan.clean()
        dev.p.upload(np.full(want.shape, 3.0))  # stale values must not leak into the result
        api.solve_noalloc(n1, n2, n3, prob.L, prob.pbc, dev.rhs, dev.p, dev.dmx1, dev.dmx2, dev.dmx3, dev.dx3)
        api.sync()
        got = dev.p.download()
        assert np.array_equal(dev.rhs.download(), prob.rhs)
        return got, want
    finally:
        dev.free()


def test_noalloc_matches_plan_solve_on_two_calls():
    first = W.smoothed_normal((128, 96, 160), (False, False, False), 1.5, seed=31)
    second = W.smoothed_normal((48, 32, 24), (True, False, False), 1.2, seed=32)
    for prob in (first, second):
        print(f"noalloc {prob.name}", flush=True)
        got, want = _both(prob)
        err = _rel(got, want)
        print(f"  rel L2 vs plan solve = {err:.3e}")
        assert err <= 1e-12


def test_noalloc_rejects_beyond_scratch_and_fp32():
    prob = W.smoothed_normal((16, 16, 8), (False, False, False), 1.5, seed=33)
    ax, ay, az = prob.ax
    dev = api.Problem3D(prob.rhs, ax, ay, az)
    try:
        with pytest.raises(api.PPFError) as e:  # 256 x 258 x 130 doubles > 64 MiB (checked before any launch)
            api.solve_noalloc(256, 256, 128, prob.L, prob.pbc, dev.rhs, dev.p, dev.dmx1, dev.dmx2, dev.dmx3, dev.dx3)
        assert e.value.code == api.PPF_ERR_SCRATCH
        d = api.Desc(16, 16, 8, 2.0, 2.0, 2.0, 0, 0, 0, api.PPF_PREC_FP32_IR, 1e-10, 8)
        rc = api.lib().ppf_solve_noalloc(ctypes.byref(d), dev.rhs.ptr, dev.p.ptr, dev.dmx1.ptr, dev.dmx2.ptr,
                                         dev.dmx3.ptr, dev.dx3.ptr, None)
        assert rc == -5  # PPF_ERR_UNSUPPORTED
    finally:
        dev.free()
