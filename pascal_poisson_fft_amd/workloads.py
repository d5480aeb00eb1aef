"""Synthetic problems for the solve path (host side, FP64, seeded).

Each builder returns a :class:`Problem` holding the grid axes and an FP64
right-hand side laid out (n3m, n2m, n1m), x fastest, ready to upload.

The operator used to build manufactured right-hand sides is the 7-point
operator the solve inverts: uniform x/y spacing L/nm with zero-flux (mirror)
Neumann walls or periodic wrap -- the operators diagonalised by the DCT-II / DFT
with the modified wavenumbers of 01_Fortran/src/fft_3d_poisson.f90:288-351 --
and the stretched z rows of fft_3d_poisson.f90:1137-1149.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .grid import Axis, cell_centres, make_axis


@dataclass
class Problem:
    name: str
    ax: tuple  # (Axis x, Axis y, Axis z)
    rhs: np.ndarray  # (n3m, n2m, n1m) float64
    exact: np.ndarray | None = None  # manufactured solution (interior), if any

    @property
    def shape(self):
        return tuple(a.nm for a in self.ax)

    @property
    def pbc(self):
        return tuple(a.pbc for a in self.ax)

    @property
    def L(self):
        return tuple(a.L for a in self.ax)


def z_rows(az: Axis):
    """Lower/upper coefficients am_k, ap_k of the z rows (fft_3d_poisson.f90:1143-1146)."""
    n = az.nm
    k = np.arange(1, n + 1)
    am = 1.0 / az.dx[k] / az.dmx[k]
    ap = 1.0 / az.dx[k] / az.dmx[k + 1]
    am[0] = 0.0
    ap[-1] = 0.0
    return am, ap


def apply_operator(p: np.ndarray, ax: Axis, ay: Axis, az: Axis) -> np.ndarray:
    """FP64 7-point operator on interior values p (n3m, n2m, n1m)."""
    out = np.zeros_like(p)
    for axis, a in ((2, ax), (1, ay)):
        h2 = (a.L / a.nm) ** 2
        if a.pbc:
            pm = np.roll(p, 1, axis=axis)
            pp = np.roll(p, -1, axis=axis)
        else:
            pm = np.concatenate([p.take([0], axis=axis), p.take(range(0, a.nm - 1), axis=axis)], axis=axis)
            pp = np.concatenate([p.take(range(1, a.nm), axis=axis), p.take([a.nm - 1], axis=axis)], axis=axis)
        out += (pp - 2.0 * p + pm) / h2
    am, ap = z_rows(az)
    zm = np.concatenate([np.zeros_like(p[:1]), p[:-1]], axis=0)
    zp = np.concatenate([p[1:], np.zeros_like(p[:1])], axis=0)
    out += am[:, None, None] * zm - (am + ap)[:, None, None] * p + ap[:, None, None] * zp
    return out


def remove_weighted_mean(f: np.ndarray, az: Axis) -> np.ndarray:
    """Subtract the dx3-weighted mean, the compatibility condition of the
    zero-flux operator on a stretched z grid (sum_ijk f_ijk dx3_k = 0)."""
    w = az.dx[1:az.nm + 1]
    m = float(np.einsum("kji,k->", f, w)) / (w.sum() * f.shape[1] * f.shape[2])
    return f - m


def _box3(f: np.ndarray, pbc) -> np.ndarray:
    """One 3x3x3 box mean: separable 3-point means, wrap on periodic axes,
    mirror (edge value repeated) at walls."""
    g = f
    for axis, periodic in ((2, pbc[0]), (1, pbc[1]), (0, pbc[2])):
        if periodic:
            g = (np.roll(g, 1, axis=axis) + g + np.roll(g, -1, axis=axis)) / 3.0
        else:
            n = g.shape[axis]
            lo = np.concatenate([g.take([0], axis=axis), g.take(range(0, n - 1), axis=axis)], axis=axis)
            hi = np.concatenate([g.take(range(1, n), axis=axis), g.take([n - 1], axis=axis)], axis=axis)
            g = (lo + g + hi) / 3.0
    return g


def manufactured_c1(n: int = 64, seed: int = 1, noise: float = 1e-3) -> Problem:
    """64^3, periodic x,y on [0,2pi]^2, Neumann z on [0,1], uniform.
    u = sin(2x)cos(3y)cos(pi z) at cell centres; f = L_h u + U(-noise, noise)
    with its global mean removed."""
    ax = make_axis(n, 2.0 * np.pi, True)
    ay = make_axis(n, 2.0 * np.pi, True)
    az = make_axis(n, 1.0, False)
    x, y, z = cell_centres(ax), cell_centres(ay), cell_centres(az)
    u = np.sin(2 * x)[None, None, :] * np.cos(3 * y)[None, :, None] * np.cos(np.pi * z)[:, None, None]
    f = apply_operator(u, ax, ay, az)
    rng = np.random.default_rng(seed)
    f = f + rng.uniform(-noise, noise, size=f.shape)
    f = f - f.mean()
    return Problem("C1 64^3 P-P uniform-z manufactured", (ax, ay, az), f, exact=u - u.mean())


def smoothed_normal(shape, pbc, gamma: float, L=(2.0, 2.0, 2.0), seed: int = 2, uniform_z: bool = False) -> Problem:
    """Seeded N(0,1) field smoothed by one 3x3x3 box mean, dx3-weighted mean removed."""
    n1, n2, n3 = shape
    ax = make_axis(n1, L[0], pbc[0])
    ay = make_axis(n2, L[1], pbc[1])
    az = make_axis(n3, L[2], pbc[2], uniform=uniform_z, gamma=gamma)
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((n3, n2, n1))
    f = _box3(f, pbc)
    f = remove_weighted_mean(f, az)
    bc = "".join("P" if p else "N" for p in pbc[:2])
    return Problem(f"{n1}x{n2}x{n3} {bc} tanh-z({gamma}) smoothed normal", (ax, ay, az), f)


def config2(n: int = 512, seed: int = 2) -> Problem:
    """512^3 FP64, Neumann x and y, tanh z (stretch 1.5) with Neumann walls, L = 2."""
    return smoothed_normal((n, n, n), (False, False, False), 1.5, seed=seed)


def config4(nprob: int = 96, seed: int = 4):
    """Many independent problems of uneven shapes: nx, ny from
    {32, 48, 64, 96, 128, 192, 256}, nz uniform in [32, 384] (even, as the
    solver requires), the first half periodic in x and y, the second half
    Neumann, z walls uniform (p = 1/2) or tanh-stretched (factor U[1, 2]),
    L = 1, smoothed N(0,1) RHS with the dx3-weighted mean removed."""
    rng = np.random.default_rng(seed)
    sizes = np.array([32, 48, 64, 96, 128, 192, 256])
    probs = []
    for i in range(nprob):
        n1, n2 = (int(v) for v in rng.choice(sizes, 2))
        n3 = 2 * int(rng.integers(16, 193))
        per = i < nprob // 2
        uniform = bool(rng.random() < 0.5)
        gamma = 0.0 if uniform else float(rng.uniform(1.0, 2.0))
        probs.append(smoothed_normal((n1, n2, n3), (per, per, False), gamma, L=(1.0, 1.0, 1.0),
                                     seed=seed * 1000 + i, uniform_z=uniform))
    return probs


def projection_tables(ax: Axis, ay: Axis, az: Axis, nmodes: int = 16, seed: int = 5, kmax: int = 24):
    """1-D mode factors of a random staggered velocity field for
    ppf_synth_divergence_rhs: nmodes separable modes with integer wavenumbers
    in [1, kmax] per axis, N(0,1) amplitudes scaled by |k|^(-5/6) (a k^-5/3
    energy spectrum), x periodic (sin/cos of 2 pi k x / L1), y and z walls
    (the wall-normal component is sin(pi k s / L), zero on both walls).
    Returns (tables[9, nmodes, stride], stride)."""
    rng = np.random.default_rng(seed)
    n1, n2, n3 = ax.nm, ay.nm, az.nm
    stride = max(n1, n2, n3) + 1
    T = np.zeros((9, nmodes, stride))
    xf = np.arange(n1) * (ax.L / n1)                      # x-faces (periodic)
    xc = (np.arange(n1) + 0.5) * (ax.L / n1)
    yf = np.arange(n2 + 1) * (ay.L / n2)                  # y-faces incl. both walls
    yc = (np.arange(n2) + 0.5) * (ay.L / n2)
    zf = az.x[1:n3 + 2] - az.x[1]                         # z-faces incl. both walls
    zc = 0.5 * (az.x[1:n3 + 1] + az.x[2:n3 + 2]) - az.x[1]
    Lz = zf[-1]
    for m in range(nmodes):
        kx, ky, kz = rng.integers(1, kmax + 1, 3)
        ph = rng.uniform(0, 2 * np.pi)
        a = rng.standard_normal(3) * float(np.sqrt(kx * kx + ky * ky + kz * kz)) ** (-5.0 / 6.0)
        T[0, m, :n1] = a[0] * np.sin(2 * np.pi * kx * xf / ax.L + ph)
        T[1, m, :n2] = np.cos(np.pi * ky * yc / ay.L)
        T[2, m, :n3] = np.cos(np.pi * kz * zc / Lz)
        T[3, m, :n1] = a[1] * np.cos(2 * np.pi * kx * xc / ax.L + ph)
        T[4, m, :n2 + 1] = np.sin(np.pi * ky * yf / ay.L)
        T[4, m, 0] = T[4, m, n2] = 0.0
        T[5, m, :n3] = np.cos(np.pi * kz * zc / Lz)
        T[6, m, :n1] = a[2] * np.cos(2 * np.pi * kx * xc / ax.L + ph)
        T[7, m, :n2] = np.cos(np.pi * ky * yc / ay.L)
        T[8, m, :n3 + 1] = np.sin(np.pi * kz * zf / Lz)
        T[8, m, 0] = T[8, m, n3] = 0.0
    return T, stride
