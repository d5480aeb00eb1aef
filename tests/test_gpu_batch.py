"""GPU checks of the batched solve (ppf_batch_*) and the z-pass autotuner.

A batch of problems with uneven extents, mixed BCs, uniform/tanh z, FP64 and
FP32_IR. The first solve runs eagerly (and learns the refinement sweeps); later
solves replay one hipGraph. Each replay is checked against the eager result and
against the oracle. Eager host work (zeroing P, synchronisation) runs between
replays.

Tolerances: FP64 problems match the oracle to <= 1e-12 relative L2. FP32_IR
problems reach the relative residual 1e-10 and are <= 1e-10 from the oracle.
A graph replay reproduces the eager result bitwise: it runs the same kernels on
the same data.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pascal_poisson_fft_amd import api
from pascal_poisson_fft_amd import workloads as W
from pascal_poisson_fft_amd.grid import make_axis

pytestmark = pytest.mark.gpu

SHAPES = [  # n1, n2, n3, pbc1, pbc2, uniform z, gamma, precision
    (32, 48, 40, True, True, True, 0.0, api.PPF_PREC_FP32_IR),
    (64, 32, 72, False, False, False, 1.4, api.PPF_PREC_FP32_IR),
    (48, 64, 33 * 2, True, False, False, 1.9, api.PPF_PREC_FP64),
    (96, 24, 128, False, True, True, 0.0, api.PPF_PREC_FP64),
    (128, 128, 32, False, False, False, 1.0, api.PPF_PREC_FP32_IR),
    (32, 48, 40, True, True, True, 0.0, api.PPF_PREC_FP32_IR),  # same shape as problem 0: shares its tuning
]


def _problems(seed0=40):
    out = []
    for i, (n1, n2, n3, p1, p2, uz, g, prec) in enumerate(SHAPES):
        L = (1.0, 1.0, 1.0)
        ax = make_axis(n1, L[0], p1)
        ay = make_axis(n2, L[1], p2)
        az = make_axis(n3, L[2], False, uniform=uz, gamma=g)
        f = W.remove_weighted_mean(np.random.default_rng(seed0 + i).standard_normal((n3, n2, n1)), az)
        dev = api.Problem3D(f, ax, ay, az)
        kw = dict(n1m=n1, n2m=n2, n3m=n3, L=L, pbc=(p1, p2, False), precision=prec, tol=1e-10, max_sweeps=8)
        ref = O.solve(f, L, (p1, p2, False), ax.dmx, ay.dmx, az.dmx, az.dx)
        out.append((kw, dev, ref))
    return out


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("flags", [0, api.PPF_BATCH_NO_GRAPH, api.PPF_BATCH_NO_TUNE])
def test_batch_eager_then_graph_replays(flags):
    probs = _problems()
    print(f"batch flags={flags}: {len(probs)} problems", flush=True)
    b = api.Batch([(kw, dev) for kw, dev, _ in probs], flags=flags)
    try:
        st = b.solve(stats=True)
        eager = [dev.p.download() for _, dev, _ in probs]
        for (kw, _, ref), P, (sw, res) in zip(probs, eager, st):
            if kw["precision"] == api.PPF_PREC_FP64:
                assert sw == 0 and _rel(P, ref) <= 1e-12
            else:
                assert 1 <= sw <= 5 and res <= 1e-10 and _rel(P, ref) <= 1e-10
        for rep in range(3):
            for _, dev, _ in probs:
                dev.p.fill_zero()
            api.sync()
            b.solve()
            api.sync()
            for (kw, dev, ref), P0 in zip(probs, eager):
                P = dev.p.download()
                assert np.array_equal(P, P0), f"replay {rep} differs from the eager solve"
        st2 = b.solve(stats=True)
        for (kw, _, _), (sw, res) in zip(probs, st2):
            if kw["precision"] == api.PPF_PREC_FP32_IR:
                assert res <= 1e-10
    finally:
        b.destroy()
        for _, dev, _ in probs:
            dev.free()


def _c4_batch_parity(nprob, seed, flags=0):
    """nprob FP64 problems drawn from the 96-problem workload's distribution
    (nx, ny in {32..256}, nz in [32, 384], half periodic x,y, half Neumann,
    uniform or tanh z) solved as one batch: each problem matches the oracle at
    1e-12 with a certified residual, and two graph replays reproduce the eager
    solve bitwise."""
    probs = W.config4(nprob=nprob, seed=seed)
    devs = [api.Problem3D(p.rhs, *p.ax) for p in probs]
    spec = [dict(n1m=p.rhs.shape[2], n2m=p.rhs.shape[1], n3m=p.rhs.shape[0], L=p.L, pbc=p.pbc) for p in probs]
    refs = [O.solve(p.rhs, p.L, p.pbc, p.ax[0].dmx, p.ax[1].dmx, p.ax[2].dmx, p.ax[2].dx) for p in probs]
    try:
        print(f"batch: {len(probs)} problems, {sum(p.rhs.size for p in probs)} cells, flags {flags}", flush=True)
        b = api.Batch(list(zip(spec, devs)), flags=flags)
        try:
            st = b.solve(stats=True)
            eager = [d.p.download() for d in devs]
            for i, (P, ref, (sw, res)) in enumerate(zip(eager, refs, st)):
                err = _rel(P, ref)
                assert err <= 1e-12 and res <= 1e-11 and sw == 0, (i, spec[i], err, res)
            for rep in range(2):
                for d in devs:
                    d.p.fill_zero()
                api.sync()
                b.solve()
                api.sync()
                for i, (d, P0) in enumerate(zip(devs, eager)):
                    assert np.array_equal(d.p.download(), P0), f"replay {rep} differs (problem {i})"
        finally:
            b.destroy()
    finally:
        for d in devs:
            d.free()


def test_batch_c4_distribution_matches_oracle():
    """24 problems of the 96-problem workload's distribution through the
    per-problem graph branches."""
    _c4_batch_parity(24, 11, api.PPF_BATCH_NO_TUNE)


@pytest.mark.skip(reason="the fused batch launch (one persistent kernel over a device work queue, "
                         "ppf_fused.hpp) did not complete in either of its two GPU runs and the cause "
                         "is not shown; the path is switched off (kFusedBatch in ppf_batch.hip)")
def test_fused_batch_matches_oracle():
    """The fused launch on 24 problems of the 96-problem workload's
    distribution, to the same bar as the per-problem branches."""
    _c4_batch_parity(24, 11)


def test_tune_picks_a_variant_and_keeps_parity():
    prob = W.smoothed_normal((64, 64, 96), (False, False, False), 1.5, seed=9)
    ax, ay, az = prob.ax
    plan = api.FFTPoisson(64, 64, 96, prob.L, prob.pbc)
    dev = api.Problem3D(prob.rhs, ax, ay, az)
    try:
        args = (dev.rhs, dev.p, dev.dmx1, dev.dmx2, dev.dmx3, dev.dx3)
        v = plan.tune(*args)
        assert v > 0
        plan.solve(*args)
        api.sync()
        ref = O.solve(prob.rhs, prob.L, prob.pbc, ax.dmx, ay.dmx, az.dmx, az.dx)
        assert _rel(dev.p.download(), ref) <= 1e-12
    finally:
        dev.free()
        plan.clean()


def test_fp32_ir_replay_on_a_rhs_needing_more_sweeps():
    """The graph is captured after an eager solve of a zero right-hand side,
    which needs no sweep.  A replay on a real right-hand side must still reach
    the tolerance (the device-side test adds the sweeps) -- or report
    PPF_ERR_NOT_CONVERGED -- never return silently above it."""
    n1, n2, n3 = 48, 32, 40
    L = (1.0, 1.0, 1.0)
    ax, ay = make_axis(n1, L[0], False), make_axis(n2, L[1], False)
    az = make_axis(n3, L[2], False, uniform=False, gamma=1.5)
    f = W.remove_weighted_mean(np.random.default_rng(5).standard_normal((n3, n2, n1)), az)
    dev = api.Problem3D(np.zeros_like(f), ax, ay, az)
    kw = dict(n1m=n1, n2m=n2, n3m=n3, L=L, pbc=(False, False, False), precision=api.PPF_PREC_FP32_IR,
              tol=1e-10, max_sweeps=8)
    b = api.Batch([(kw, dev)])
    try:
        (sw0, _), = b.solve(stats=True)
        assert sw0 == 0
        dev.rhs.upload(f)
        b.solve()
        (sw, res), = b.solve(stats=True)
        P = dev.p.download()
        ref = O.solve(f, L, (False, False, False), ax.dmx, ay.dmx, az.dmx, az.dx)
        print(f"learned sweeps={sw0}; replay sweeps={sw} res={res:.3e} err={_rel(P, ref):.3e}")
        assert sw >= 1 and res <= 1e-10 and _rel(P, ref) <= 1e-10
    finally:
        b.destroy()
        dev.free()
