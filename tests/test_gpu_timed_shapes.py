"""GPU parity at the shapes the benchmark times, against the FP64 oracle.

The benchmarked kernels are template instantiations per line length, so the
512- and 1024-point x/y passes, the row-blocked y mapping at 512 rows, the
register-capped refinement forms (lines of 512 points and more) and the z pass
at 512 rows are only checked if a test runs exactly those lengths.

Tolerances: FP64 solves match the oracle to <= 1e-12 relative L2 of the full
ghosted solution (the drop-in target).  FP32_IR solves reach the relative
residual 1e-10 and match the oracle to <= 1e-10 (measured: about 1e-11).
Full-size 1024x1024x768 solves, too large for the oracle in a test, are held
to size-independent properties: the FP64 residual of the 7-point operator
(<= 1e-12) and, for FP32_IR graph replays, the residual tolerance on a fresh
right-hand side.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pascal_poisson_fft_amd import api
from pascal_poisson_fft_amd import workloads as W
from pascal_poisson_fft_amd.grid import make_axis

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_IR_RES = 1e-10
TOL_IR_ORACLE = 1e-10


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _solve(prob, precision=api.PPF_PREC_FP64, **kw):
    ax, ay, az = prob.ax
    n1, n2, n3 = prob.shape
    plan = api.FFTPoisson(n1, n2, n3, prob.L, prob.pbc, precision=precision, **kw)
    dev = api.Problem3D(prob.rhs, ax, ay, az)
    try:
        st = plan.solve(dev.rhs, dev.p, dev.dmx1, dev.dmx2, dev.dmx3, dev.dx3)
        api.sync()
        res = plan.residual(dev.rhs, dev.p, dev.dmx3, dev.dx3)
        return dev.p.download(), st, res
    finally:
        dev.free()
        plan.clean()


def _oracle(prob):
    ax, ay, az = prob.ax
    return O.solve(prob.rhs, prob.L, prob.pbc, ax.dmx, ay.dmx, az.dmx, az.dx)


@pytest.mark.parametrize("pbc", [(False, False, False), (True, False, False), (True, True, False),
                                 (False, True, False)], ids=["NN", "PN", "PP", "NP"])
def test_512_lines_fp64(pbc):
    """512-point x and y lines (the timed instantiations) with 64 z rows."""
    print(f"512x512x64 {pbc} FP64", flush=True)
    prob = W.smoothed_normal((512, 512, 64), pbc, 1.5, seed=21)
    got, _, res = _solve(prob)
    err = _rel(got, _oracle(prob))
    print(f"  err_vs_oracle={err:.3e} residual={res:.3e}")
    assert err <= TOL_FP64 and res <= 1e-12


def test_512_cube_fp64_headline():
    """The headline workload itself: 512^3, Neumann x and y, tanh z 1.5, smoothed N(0,1) RHS."""
    print("512^3 N-N tanh 1.5 FP64", flush=True)
    prob = W.config2(512)
    got, _, res = _solve(prob)
    err = _rel(got, _oracle(prob))
    print(f"  err_vs_oracle={err:.3e} residual={res:.3e}")
    assert err <= TOL_FP64 and res <= 1e-12


@pytest.mark.parametrize("pbc", [(False, False, False), (True, False, False)], ids=["NN", "PN"])
def test_512_lines_fp32_ir(pbc):
    """The register-capped refinement forms of passes 1 and 5 (lines >= 512)."""
    print(f"512x512x64 {pbc} FP32_IR", flush=True)
    prob = W.smoothed_normal((512, 512, 64), pbc, 1.5, seed=22)
    got, st, res = _solve(prob, api.PPF_PREC_FP32_IR, tol=TOL_IR_RES, max_sweeps=8)
    err = _rel(got, _oracle(prob))
    print(f"  sweeps={st.sweeps} res={st.rel_residual:.3e} recomputed={res:.3e} err_vs_oracle={err:.3e}")
    assert st.rel_residual <= TOL_IR_RES and res <= 1.05 * TOL_IR_RES
    assert err <= TOL_IR_ORACLE


def _projection_problem(n1, n2, n3, gamma, seed, L=(2.0, 2.0, 2.0)):
    """The projection-stream right-hand side (divergence of a random staggered
    velocity field), generated on the GPU by ppf_synth_divergence_rhs."""
    ax, ay = make_axis(n1, L[0], True), make_axis(n2, L[1], False)
    az = make_axis(n3, L[2], False, uniform=False, gamma=gamma)
    T, stride = W.projection_tables(ax, ay, az, nmodes=16, seed=seed)
    f = api.DeviceArray((n3, n2, n1))
    dx3 = api.DeviceArray.from_host(az.dx)
    tabs = api.DeviceArray.from_host(np.ascontiguousarray(T))
    try:
        api.synth_divergence_rhs(f, n1, n2, n3, L[0], L[1], dx3, tabs, 16, stride, seed, 0.1)
        api.sync()
        rhs = f.download()
    finally:
        for a in (f, dx3, tabs):
            a.free()
    return W.Problem(f"{n1}x{n2}x{n3} P-N tanh {gamma} projection RHS", (ax, ay, az), rhs)


@pytest.mark.parametrize("precision", [api.PPF_PREC_FP64, api.PPF_PREC_FP32_IR], ids=["fp64", "fp32ir"])
def test_1024_periodic_x_projection_rhs(precision):
    """1024-point periodic x (R2C) and 1024-row Neumann y with the projection
    stream's divergence RHS and tanh z 1.8, at 32 z rows (oracle-sized)."""
    print(f"1024x1024x32 P-N projection RHS precision={precision}", flush=True)
    prob = _projection_problem(1024, 1024, 32, 1.8, seed=5)
    kw = dict(tol=TOL_IR_RES, max_sweeps=8) if precision == api.PPF_PREC_FP32_IR else {}
    got, st, res = _solve(prob, precision, **kw)
    err = _rel(got, _oracle(prob))
    print(f"  sweeps={st.sweeps} residual={res:.3e} err_vs_oracle={err:.3e}")
    if precision == api.PPF_PREC_FP64:
        assert err <= TOL_FP64 and res <= 1e-12
    else:
        assert res <= 1.05 * TOL_IR_RES and err <= TOL_IR_ORACLE


def test_projection_stream_full_size_properties():
    """1024x1024x768, periodic x, Neumann y, tanh z 1.8 (the projection
    stream): the tuned FP64 solve's residual, and an FP32_IR graph replay on a
    fresh right-hand side (the sweep count was learned on the first one)."""
    n1, n2, n3 = 1024, 1024, 768
    L = (2.0, 2.0, 2.0)
    ax, ay = make_axis(n1, L[0], True), make_axis(n2, L[1], False)
    az = make_axis(n3, L[2], False, uniform=False, gamma=1.8)
    rhs = api.DeviceArray((n3, n2, n1))
    P = api.DeviceArray((n3 + 2, n2 + 2, n1 + 2))
    P.fill_zero()
    dmx1, dmx2 = api.DeviceArray.from_host(ax.dmx), api.DeviceArray.from_host(ay.dmx)
    dmx3, dx3 = api.DeviceArray.from_host(az.dmx), api.DeviceArray.from_host(az.dx)
    tabs = None
    bufs = [rhs, P, dmx1, dmx2, dmx3, dx3]

    def gen(seed):
        nonlocal tabs
        T, stride = W.projection_tables(ax, ay, az, nmodes=16, seed=seed)
        if tabs is None:
            tabs = api.DeviceArray(T.shape)
            bufs.append(tabs)
        tabs.upload(np.ascontiguousarray(T))
        api.synth_divergence_rhs(rhs, n1, n2, n3, L[0], L[1], dx3, tabs, 16, stride, seed, 0.1)

    class _Dev:
        pass

    dev = _Dev()
    dev.rhs, dev.p, dev.dmx1, dev.dmx2, dev.dmx3, dev.dx3 = rhs, P, dmx1, dmx2, dmx3, dx3
    try:
        print("1024x1024x768 FP64 tuned solve", flush=True)
        plan = api.FFTPoisson(n1, n2, n3, L, (True, False, False))
        args = (rhs, P, dmx1, dmx2, dmx3, dx3)
        gen(5)
        v = plan.tune(*args)
        plan.solve(*args)
        api.sync()
        res64 = plan.residual(rhs, P, dmx3, dx3)
        plan.clean()
        print(f"  zvariant={v} residual={res64:.3e}", flush=True)
        assert res64 <= 1e-12
        print("1024x1024x768 FP32_IR graph replay on a fresh RHS", flush=True)
        spec = dict(n1m=n1, n2m=n2, n3m=n3, L=L, pbc=(True, False, False), precision=api.PPF_PREC_FP32_IR,
                    tol=TOL_IR_RES, max_sweeps=8)
        b = api.Batch([(spec, dev)])
        try:
            gen(5)
            (sw0, r0), = b.solve(stats=True)  # eager: learns the sweeps, captures the graph
            gen(6)
            api.sync()
            b.solve()
            api.sync()
            (sw1, r1), = b.solve(stats=True)
            print(f"  learned sweeps={sw0} res={r0:.3e}; replay on fresh RHS sweeps={sw1} res={r1:.3e}")
            assert r0 <= TOL_IR_RES and r1 <= TOL_IR_RES
        finally:
            b.destroy()
    finally:
        for a in bufs:
            a.free()
