"""GPU parity: the HIP solve path (through the C ABI) against the FP64 oracle.

Tolerance (FP64 path): relative L2 of the full ghosted solution P against the
oracle <= 1e-12, per the plain-method drop-in target; the FP64 residual of the
GPU solution <= 1e-11.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pascal_poisson_fft_amd import api
from pascal_poisson_fft_amd import workloads as W
from pascal_poisson_fft_amd.grid import make_axis

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


def _rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _problem(n1, n2, n3, pbc, uniform_z, gamma, seed, L=(2.0, 3.0, 1.5), compatible=True):
    ax = make_axis(n1, L[0], pbc[0])
    ay = make_axis(n2, L[1], pbc[1])
    az = make_axis(n3, L[2], False, uniform=uniform_z, gamma=gamma)
    f = np.random.default_rng(seed).standard_normal((n3, n2, n1))
    if compatible:
        f = W.remove_weighted_mean(f, az)
    return ax, ay, az, f


BCS = [(False, False), (True, True), (True, False), (False, True)]


@pytest.mark.parametrize("pbc", BCS, ids=["NN", "PP", "PN", "NP"])
@pytest.mark.parametrize("n1,n2,n3", [(8, 8, 8), (16, 12, 10), (32, 48, 20), (64, 64, 64), (96, 24, 6), (12, 192, 34),
                                      (8, 1024, 6), (16, 768, 4)])
def test_solve_matches_oracle(pbc, n1, n2, n3):
    ax, ay, az, f = _problem(n1, n2, n3, pbc, False, 1.5, n1 * 7 + n2 * 3 + n3)
    ref = O.solve(f, (ax.L, ay.L, az.L), (pbc[0], pbc[1], False), ax.dmx, ay.dmx, az.dmx, az.dx)
    got = api.solve_host(f, ax, ay, az)
    assert got.shape == ref.shape
    assert _rel_l2(got, ref) <= TOL_FP64, _rel_l2(got, ref)


@pytest.mark.parametrize("pbc", BCS, ids=["NN", "PP", "PN", "NP"])
@pytest.mark.parametrize("n1,n2,n3", [(160, 320, 8), (320, 160, 6), (10, 14, 12), (30, 60, 10), (640, 20, 4),
                                      (14, 896, 4), (960, 30, 4), (112, 224, 6), (480, 120, 4), (40, 80, 16)])
def test_radix5_7_lengths_match_oracle(pbc, n1, n2, n3):
    """Lengths with factors 5, 7 and 15 (radix-5 and radix-7 Stockham stages),
    which the reference's cuFFT plans accept: 5*2^a (10..640), 7*2^a (14..896),
    15*2^a (30..960)."""
    ax, ay, az, f = _problem(n1, n2, n3, pbc, False, 1.5, n1 * 5 + n2 * 7 + n3)
    ref = O.solve(f, (ax.L, ay.L, az.L), (pbc[0], pbc[1], False), ax.dmx, ay.dmx, az.dmx, az.dx)
    got = api.solve_host(f, ax, ay, az)
    assert _rel_l2(got, ref) <= TOL_FP64, _rel_l2(got, ref)


@pytest.mark.parametrize("pbc", BCS, ids=["NN", "PP", "PN", "NP"])
def test_incompatible_rhs_pin_semantics(pbc):
    """RHS without the compatibility projection: the (0,0)-mode pin decides the
    answer; it must match the oracle exactly as the reference's pin does."""
    ax, ay, az, f = _problem(24, 16, 18, pbc, True, 0.0, 11, compatible=False)
    f = f + 0.3
    ref = O.solve(f, (ax.L, ay.L, az.L), (pbc[0], pbc[1], False), ax.dmx, ay.dmx, az.dmx, az.dx)
    got = api.solve_host(f, ax, ay, az)
    assert _rel_l2(got, ref) <= TOL_FP64


@pytest.mark.parametrize("n", [128, 256])
def test_solve_matches_oracle_large(n):
    prob = W.smoothed_normal((n, n, n // 2), (False, False, False), 1.5, seed=5)
    ax, ay, az = prob.ax
    ref = O.solve(prob.rhs, prob.L, prob.pbc, ax.dmx, ay.dmx, az.dmx, az.dx)
    got = api.solve_host(prob.rhs, ax, ay, az)
    assert _rel_l2(got, ref) <= TOL_FP64


def test_c1_manufactured_64():
    prob = W.manufactured_c1(64)
    ax, ay, az = prob.ax
    got = api.solve_host(prob.rhs, ax, ay, az)
    ref = O.solve(prob.rhs, prob.L, prob.pbc, ax.dmx, ay.dmx, az.dmx, az.dx)
    assert _rel_l2(got, ref) <= TOL_FP64
    assert np.abs(got[1:-1, 1:-1, 1:-1] - prob.exact).max() < 1e-3
    assert O.rel_residual(prob.rhs, got, prob.L, prob.pbc, az.dmx, az.dx) < 1e-11


@pytest.mark.parametrize("pbc", BCS, ids=["NN", "PP", "PN", "NP"])
@pytest.mark.parametrize("n1,n2,n3", [(32, 24, 16), (96, 20, 40), (320, 10, 34), (160, 28, 6), (64, 12, 70)])
def test_gpu_residual_matches_oracle(pbc, n1, n2, n3):
    """The residual kernel on a random iterate: y blocks of 8, 4 and 2 rows
    (n2m = 24; 20, 12, 28; 10 -- extents are even, so never 1), one and two x
    tiles with the last cell inside a wave (n1m = 32, 160), z blocks shorter
    than 32 planes, every x/y branch.
    A wrong neighbour in any row or column of cells moves the sum of r^2 by
    far more than the 1e-12 bound."""
    ax, ay, az, f = _problem(n1, n2, n3, pbc, False, 1.2, 3)
    P = np.random.default_rng(4).standard_normal((n3 + 2, n2 + 2, n1 + 2))
    plan = api.FFTPoisson(n1, n2, n3, (ax.L, ay.L, az.L), (pbc[0], pbc[1], False))
    fd, pd = api.DeviceArray.from_host(f), api.DeviceArray.from_host(P)
    d3, x3 = api.DeviceArray.from_host(az.dmx), api.DeviceArray.from_host(az.dx)
    r = plan.residual(fd, pd, d3, x3)
    ref = O.rel_residual(f, P, (ax.L, ay.L, az.L), pbc, az.dmx, az.dx)
    assert abs(r - ref) <= 1e-12 * ref, (r, ref)
    plan.clean()


@pytest.mark.parametrize("pbc", [(False, False), (True, True)], ids=["NN", "PP"])
@pytest.mark.parametrize("n1,n2,n3", [(512, 256, 512), (512, 256, 256), (512, 256, 128)])
def test_gpu_residual_row_blocks_match_oracle(pbc, n1, n2, n3):
    """Shapes large enough that the residual kernel blocks 8, 4 and 2 rows per
    thread (it keeps at least 4 tiles per CU on a 256-CU device: 512 x 256 x
    512, 256, 128 cells give 8, 4, 2)."""
    ax, ay, az, f = _problem(n1, n2, n3, pbc, False, 1.2, 5)
    P = np.random.default_rng(6).standard_normal((n3 + 2, n2 + 2, n1 + 2))
    plan = api.FFTPoisson(n1, n2, n3, (ax.L, ay.L, az.L), (pbc[0], pbc[1], False))
    fd, pd = api.DeviceArray.from_host(f), api.DeviceArray.from_host(P)
    d3, x3 = api.DeviceArray.from_host(az.dmx), api.DeviceArray.from_host(az.dx)
    try:
        r = plan.residual(fd, pd, d3, x3)
    finally:
        plan.clean()
        for a in (fd, pd, d3, x3):
            a.free()
    ref = O.rel_residual(f, P, (ax.L, ay.L, az.L), pbc, az.dmx, az.dx)
    assert abs(r - ref) <= 1e-12 * ref, (r, ref)


def test_repeated_solves_are_deterministic_and_leave_rhs_intact():
    ax, ay, az, f = _problem(64, 32, 40, (False, False), False, 1.5, 8)
    plan = api.FFTPoisson(64, 32, 40, (ax.L, ay.L, az.L), (False, False, False))
    prob = api.Problem3D(f, ax, ay, az)
    plan.solve(prob.rhs, prob.p, prob.dmx1, prob.dmx2, prob.dmx3, prob.dx3)
    api.sync()
    a = prob.p.download()
    prob.p.upload(np.full(a.shape, 7.0))  # stale ghosts/values must not leak in
    plan.solve(prob.rhs, prob.p, prob.dmx1, prob.dmx2, prob.dmx3, prob.dx3)
    api.sync()
    b = prob.p.download()
    assert np.array_equal(a, b)
    assert np.array_equal(prob.rhs.download(), f)
    prob.free()
    plan.clean()


@pytest.mark.parametrize("nx,ny,nz", [(32, 8, 64), (7, 3, 5), (100, 1, 2), (512, 4, 512)])
def test_tdma_many_bitwise_vs_reference_order(nx, ny, nz):
    rng = np.random.default_rng(nx + ny + nz)
    a = rng.uniform(0.1, 1.0, (nz, ny, nx))
    c = rng.uniform(0.1, 1.0, (nz, ny, nx))
    b = -(a + c) - rng.uniform(0.0, 2.0, (nz, ny, nx))
    a[0] = 0.0
    c[-1] = 0.0
    d = rng.standard_normal((nz, ny, nx))
    ref = O.tdma_many(a.reshape(nz, -1), b.reshape(nz, -1), c.reshape(nz, -1), d.reshape(nz, -1)).reshape(d.shape)
    if O.ref_tdma_available():
        ref2 = O.ref_tdma_many(a.reshape(nz, -1), b.reshape(nz, -1), c.reshape(nz, -1), d.reshape(nz, -1))
        assert np.array_equal(ref, ref2.reshape(d.shape))
    plan = api.TDMAPlanMany(nx, ny, nz)
    A, B, Cd, D = (api.DeviceArray.from_host(x) for x in (a, b, c, d))
    plan.solve(A, B, Cd, D)
    api.sync()
    got = D.download()
    assert np.array_equal(got, ref)
    plan.destroy()


def test_plan_rejects_bad_sizes():
    with pytest.raises(api.PPFError) as e:
        api.FFTPoisson(22, 8, 8)  # 22 = 2*11: no compiled line length
    assert e.value.code == -2
    with pytest.raises(api.PPFError) as e:
        api.FFTPoisson(8, 1280, 8)  # beyond 1024
    assert e.value.code == -2
    with pytest.raises(api.PPFError) as e:
        api.FFTPoisson(8, 8, 7)  # odd z
    assert e.value.code == -2
    with pytest.raises(api.PPFError) as e:
        api.FFTPoisson(8, 8, 4096 + 2)  # z lines beyond the supported 4096 rows
    assert e.value.code == -2


@pytest.mark.parametrize("pbc", BCS, ids=["NN", "PP", "PN", "NP"])
@pytest.mark.parametrize("n1,n2,n3", [(16, 12, 10), (64, 32, 48)])
def test_periodic_z_matches_oracle(pbc, n1, n2, n3):
    """pbc3: the reference's z rows (walls' am/ap zeroing kept) with wrapped z
    ghost planes, on a periodic z grid."""
    L = (2.0, 3.0, 1.5)
    ax = make_axis(n1, L[0], pbc[0])
    ay = make_axis(n2, L[1], pbc[1])
    az = make_axis(n3, L[2], True)
    f = np.random.default_rng(n1 + n2 + n3).standard_normal((n3, n2, n1))
    f -= f.mean()
    got = api.solve_host(f, ax, ay, az)
    ref = O.solve(f, L, (pbc[0], pbc[1], True), ax.dmx, ay.dmx, az.dmx, az.dx)
    assert _rel_l2(got, ref) <= TOL_FP64


@pytest.mark.parametrize("nx,ny,nz", [(32, 8, 64), (7, 3, 5), (100, 1, 2), (512, 4, 512)])
def test_tdma_cycl_many_bitwise_vs_reference_order(nx, ny, nz):
    rng = np.random.default_rng(3 * nx + ny + nz)
    a = rng.uniform(0.1, 1.0, (nz, ny, nx))
    c = rng.uniform(0.1, 1.0, (nz, ny, nx))
    b = -(a + c) - rng.uniform(0.5, 2.0, (nz, ny, nx))
    d = rng.standard_normal((nz, ny, nx))
    flat = [x.reshape(nz, -1) for x in (a, b, c, d)]
    ref = O.tdma_cycl_many(*flat).reshape(d.shape)
    plan = api.TDMAPlanMany(nx, ny, nz)
    A, B, Cd, D = (api.DeviceArray.from_host(x) for x in (a, b, c, d))
    plan.solve_cycle(A, B, Cd, D)
    api.sync()
    assert np.array_equal(D.download(), ref)
    plan.destroy()
