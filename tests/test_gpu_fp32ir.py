"""GPU checks of the FP32 + iterative-refinement precision (PPF_PREC_FP32_IR).

The transforms and the Thomas storage are FP32.  The residual f - L x, the
correction x += e, the pinned (0,0) z system and the mean stay FP64.  Pass
criteria:
  * the solve stops with ||f - L x||/||f|| <= tol = 1e-10. The library's own
    FP64 residual kernel recomputes it independently (<= 1.05e-10);
  * relative L2 of the full ghosted P against the FP64 oracle <= 1e-10 (the
    measured values are about 1e-12 to 1e-11 and are printed);
  * the sweep count stays small (<= 5), as contraction by an FP32 solve implies.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pascal_poisson_fft_amd import api
from pascal_poisson_fft_amd import workloads as W
from pascal_poisson_fft_amd.grid import make_axis

pytestmark = pytest.mark.gpu

TOL = 1e-10
ERR_VS_ORACLE = 1e-10


def _rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _setup(n1, n2, n3, pbc, uniform_z, gamma, seed, L=(2.0, 3.0, 1.5)):
    ax = make_axis(n1, L[0], pbc[0])
    ay = make_axis(n2, L[1], pbc[1])
    az = make_axis(n3, L[2], False, uniform=uniform_z, gamma=gamma)
    f = np.random.default_rng(seed).standard_normal((n3, n2, n1))
    f = W.remove_weighted_mean(f, az)
    return ax, ay, az, f


def _run(ax, ay, az, f, **kw):
    n3, n2, n1 = f.shape
    plan = api.FFTPoisson(n1, n2, n3, (ax.L, ay.L, az.L), (ax.pbc, ay.pbc, False),
                          precision=api.PPF_PREC_FP32_IR, **kw)
    dev = api.Problem3D(f, ax, ay, az)
    try:
        st = plan.solve(dev.rhs, dev.p, dev.dmx1, dev.dmx2, dev.dmx3, dev.dx3)
        api.sync()
        res = plan.residual(dev.rhs, dev.p, dev.dmx3, dev.dx3)
        return dev.p.download(), st, res
    finally:
        dev.free()
        plan.clean()


BCS = [(False, False), (True, True), (True, False), (False, True)]


@pytest.mark.parametrize("pbc", BCS, ids=["NN", "PP", "PN", "NP"])
@pytest.mark.parametrize("n1,n2,n3,uniform_z", [(32, 24, 16, False), (64, 64, 64, False), (96, 48, 40, True),
                                                 (128, 192, 34, False), (160, 112, 20, False)])
def test_fp32_ir_reaches_tolerance_and_matches_oracle(pbc, n1, n2, n3, uniform_z):
    ax, ay, az, f = _setup(n1, n2, n3, pbc, uniform_z, 1.5, seed=n1 + 5 * n2 + 11 * n3)
    print(f"fp32_ir {n1}x{n2}x{n3} pbc={pbc}", flush=True)
    P, st, res = _run(ax, ay, az, f, tol=TOL, max_sweeps=8)
    ref = O.solve(f, (ax.L, ay.L, az.L), (ax.pbc, ay.pbc, False), ax.dmx, ay.dmx, az.dmx, az.dx)
    err = _rel_l2(P, ref)
    print(f"  sweeps={st.sweeps} rel_res={st.rel_residual:.3e} recomputed={res:.3e} err_vs_oracle={err:.3e}")
    assert st.rel_residual <= TOL
    assert res <= 1.05 * TOL
    assert st.sweeps <= 5
    assert err <= ERR_VS_ORACLE


def test_fp32_ir_fixed_sweeps_mode():
    """tol < 0: exactly max_sweeps corrections, no convergence test."""
    ax, ay, az, f = _setup(48, 32, 24, (False, False), False, 1.5, seed=3)
    P3, st3, res3 = _run(ax, ay, az, f, tol=-1.0, max_sweeps=3)
    assert st3.sweeps == 3
    assert res3 == pytest.approx(st3.rel_residual, rel=1e-6)
    P0, st0, res0 = _run(ax, ay, az, f, tol=-1.0, max_sweeps=0)  # <= 0: the default cap 8
    assert st0.sweeps == 8
    # both are at the FP64 rounding floor of the residual, so only bound them
    assert res3 < 1e-12 and res0 < 1e-12


def test_fp32_ir_reports_not_converged():
    """A tolerance below FP64 reach stops at max_sweeps with PPF_ERR_NOT_CONVERGED."""
    ax, ay, az, f = _setup(32, 32, 16, (True, False), False, 1.2, seed=4)
    n3, n2, n1 = f.shape
    plan = api.FFTPoisson(n1, n2, n3, (ax.L, ay.L, az.L), (ax.pbc, ay.pbc, False),
                          precision=api.PPF_PREC_FP32_IR, tol=1e-30, max_sweeps=2)
    dev = api.Problem3D(f, ax, ay, az)
    try:
        with pytest.raises(api.PPFError) as ei:
            plan.solve(dev.rhs, dev.p, dev.dmx1, dev.dmx2, dev.dmx3, dev.dx3)
        assert ei.value.code == api.PPF_ERR_NOT_CONVERGED
        st = plan.solve(dev.rhs, dev.p, dev.dmx1, dev.dmx2, dev.dmx3, dev.dx3, allow_not_converged=True)
        assert st.sweeps == 2 and 0 < st.rel_residual < 1e-9
    finally:
        dev.free()
        plan.clean()


def test_fp32_ir_config_shape_256():
    """The 512^3 precision workload's structure at 256^3 (N-N, tanh z 1.5, smoothed RHS)."""
    prob = W.smoothed_normal((256, 256, 256), (False, False, False), 1.5, seed=2)
    ax, ay, az = prob.ax
    P, st, res = _run(ax, ay, az, prob.rhs, tol=TOL, max_sweeps=8)
    ref = O.solve(prob.rhs, prob.L, prob.pbc, ax.dmx, ay.dmx, az.dmx, az.dx)
    err = _rel_l2(P, ref)
    print(f"256^3 fp32_ir sweeps={st.sweeps} rel_res={st.rel_residual:.3e} err_vs_oracle={err:.3e}")
    assert st.rel_residual <= TOL and res <= 1.05 * TOL
    assert err <= ERR_VS_ORACLE


def _plan_dev(ax, ay, az, f, **kw):
    n3, n2, n1 = f.shape
    plan = api.FFTPoisson(n1, n2, n3, (ax.L, ay.L, az.L), (ax.pbc, ay.pbc, False),
                          precision=api.PPF_PREC_FP32_IR, **kw)
    return plan, api.Problem3D(f, ax, ay, az)


@pytest.mark.parametrize("pbc", [(False, False), (True, False)], ids=["NN", "PN"])
def test_fp32_ir_device_side_convergence(pbc):
    """ppf_solve_async with min_sweeps = 0: the device-side residual test alone
    decides the sweeps; the result matches the host-tested solve."""
    ax, ay, az, f = _setup(64, 48, 40, pbc, False, 1.5, seed=12)
    P_host, st_host, _ = _run(ax, ay, az, f, tol=TOL, max_sweeps=8)
    plan, dev = _plan_dev(ax, ay, az, f, tol=TOL, max_sweeps=8)
    try:
        plan.solve_async(dev.rhs, dev.p, dev.dmx1, dev.dmx2, dev.dmx3, dev.dx3, min_sweeps=0)
        st = plan.async_result()
        res = plan.residual(dev.rhs, dev.p, dev.dmx3, dev.dx3)
        P = dev.p.download()
        print(f"async sweeps={st.sweeps} res={st.rel_residual:.3e} recomputed={res:.3e}; host-tested "
              f"sweeps={st_host.sweeps}")
        assert st.sweeps == st_host.sweeps >= 1
        assert st.rel_residual <= TOL and res <= 1.05 * TOL
        assert np.array_equal(P, P_host)
        # more unconditional sweeps than needed: the device test runs none extra
        plan.solve_async(dev.rhs, dev.p, dev.dmx1, dev.dmx2, dev.dmx3, dev.dx3, min_sweeps=st.sweeps + 1)
        st2 = plan.async_result()
        assert st2.sweeps == st.sweeps + 1 and st2.rel_residual <= TOL
        # no guarded sweeps allowed: the FP32 solve alone, reported as a miss
        plan.solve_async(dev.rhs, dev.p, dev.dmx1, dev.dmx2, dev.dmx3, dev.dx3, min_sweeps=0, extra_sweeps=0)
        with pytest.raises(api.PPFError) as ei:
            plan.async_result()
        assert ei.value.code == api.PPF_ERR_NOT_CONVERGED
        st3 = plan.async_result(allow_not_converged=True)
        assert st3.sweeps == 0 and st3.rel_residual > TOL
    finally:
        dev.free()
        plan.clean()


def test_fp32_ir_device_side_not_converged():
    """A tolerance below FP64 reach: every sweep runs and the result reports the miss."""
    ax, ay, az, f = _setup(32, 32, 16, (True, False), False, 1.2, seed=4)
    plan, dev = _plan_dev(ax, ay, az, f, tol=1e-30, max_sweeps=3)
    try:
        plan.solve_async(dev.rhs, dev.p, dev.dmx1, dev.dmx2, dev.dmx3, dev.dx3, min_sweeps=0)
        with pytest.raises(api.PPFError) as ei:
            plan.async_result()
        assert ei.value.code == api.PPF_ERR_NOT_CONVERGED
        st = plan.async_result(allow_not_converged=True)
        assert st.sweeps == 3 and 0 < st.rel_residual < 1e-9
    finally:
        dev.free()
        plan.clean()
