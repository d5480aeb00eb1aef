"""GPU check of the synthetic projection-step RHS (ppf_synth_divergence_rhs).

The kernel is compared with a numpy evaluation of the same discrete
divergence, including the counter-based face noise, at 1e-12 relative. The
sum of f times the cell volumes (uniform in x, y; dx3 in z) must vanish to
rounding: the RHS is compatible with the zero-flux operator by construction.
"""
import numpy as np
import pytest

from pascal_poisson_fft_amd import api
from pascal_poisson_fft_amd import workloads as W
from pascal_poisson_fft_amd.grid import make_axis

pytestmark = pytest.mark.gpu

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _hash_unit(x):
    x = x + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    x = x ^ (x >> np.uint64(31))
    return (x >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def _key(seed, comp, i, j, k):
    return ((np.uint64(seed) * np.uint64(0x100000001B3)) ^ (np.uint64(comp) << np.uint64(62))
            ^ (k.astype(np.uint64) << np.uint64(40)) ^ (j.astype(np.uint64) << np.uint64(20)) ^ i.astype(np.uint64))


def _reference(T, ax, ay, az, seed, noise):
    n1, n2, n3 = ax.nm, ay.nm, az.nm
    k, j, i = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    ip = (i + 1) % n1
    du = np.einsum("mi,mj,mk->kji", T[0][:, (np.arange(n1) + 1) % n1] - T[0][:, :n1], T[1][:, :n2], T[2][:, :n3])
    dv = np.einsum("mi,mj,mk->kji", T[3][:, :n1], T[4][:, 1:n2 + 1] - T[4][:, :n2], T[5][:, :n3])
    dw = np.einsum("mi,mj,mk->kji", T[6][:, :n1], T[7][:, :n2], T[8][:, 1:n3 + 1] - T[8][:, :n3])
    nu0, nu1 = _hash_unit(_key(seed, 0, i, j, k)), _hash_unit(_key(seed, 0, ip, j, k))
    nv0 = np.where(j == 0, 0.0, _hash_unit(_key(seed, 1, i, j, k)))
    nv1 = np.where(j + 1 == n2, 0.0, _hash_unit(_key(seed, 1, i, j + 1, k)))
    nw0 = np.where(k == 0, 0.0, _hash_unit(_key(seed, 2, i, j, k)))
    nw1 = np.where(k + 1 == n3, 0.0, _hash_unit(_key(seed, 2, i, j, k + 1)))
    du = du + noise * (nu1 - nu0)
    dv = dv + noise * (nv1 - nv0)
    dw = dw + noise * (nw1 - nw0)
    return du * (n1 / ax.L) + dv * (n2 / ay.L) + dw / az.dx[1:n3 + 1][:, None, None]


@pytest.mark.parametrize("n1,n2,n3,nmodes", [(16, 12, 10, 3), (32, 24, 40, 8)])
def test_synth_divergence_matches_numpy_and_is_compatible(n1, n2, n3, nmodes):
    ax = make_axis(n1, 2.0, True)
    ay = make_axis(n2, 1.5, False)
    az = make_axis(n3, 1.0, False, uniform=False, gamma=1.8)
    T, stride = W.projection_tables(ax, ay, az, nmodes=nmodes, seed=n1 + n3)
    seed, noise = 77, 0.25
    ref = _reference(T, ax, ay, az, seed, noise)
    f = api.DeviceArray((n3, n2, n1))
    tab = api.DeviceArray.from_host(np.ascontiguousarray(T))
    dx3 = api.DeviceArray.from_host(az.dx)
    try:
        api.synth_divergence_rhs(f, n1, n2, n3, ax.L, ay.L, dx3, tab, nmodes, stride, seed, noise)
        api.sync()
        got = f.download()
    finally:
        for a in (f, tab, dx3):
            a.free()
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    w = az.dx[1:n3 + 1][:, None, None]
    assert abs(float((got * w).sum())) <= 1e-12 * float((np.abs(got) * w).sum())
