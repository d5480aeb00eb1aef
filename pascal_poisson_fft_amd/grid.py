"""Host-side grid metrics for one rank.

Restates the caller-side setup that feeds the solve path:
``calculate_x`` / ``non_uniform_channel_grid`` / ``uniform_grid`` /
``calculate_dx_dmx`` in 01_Fortran/examples/mpi_subdomain.f90:397-472, with the
one-rank ghost exchange ``mpi_subdomain_ghostcell_1d`` (:783-798).

Every array has nm+2 entries and element q is the reference's Fortran element q
(declared ``(0:nsub)`` with nsub = nm+1).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Axis:
    """Nodes x (cell faces), cell widths dx and centre spacings dmx of one axis."""

    x: np.ndarray
    dx: np.ndarray
    dmx: np.ndarray
    nm: int
    L: float
    pbc: bool
    uniform: bool
    gamma: float


def make_axis(nm: int, L: float, pbc: bool, uniform: bool = True, gamma: float = 0.0) -> Axis:
    """x, dx, dmx for one axis (mpi_subdomain.f90:397-472)."""
    if nm < 2:
        raise ValueError("an axis needs at least 2 cells")
    idx = np.arange(nm + 2, dtype=np.float64)
    if uniform:
        x = (idx - 1.0) * L / nm
    else:
        x = L * 0.5 * (1.0 + np.tanh(0.5 * gamma * (2.0 * (idx - 1.0) / nm - 1.0)) / np.tanh(gamma * 0.5))
    if not pbc:
        x[0] = x[1]  # wall: the ghost node sits on the wall (:413)
    dx = np.zeros(nm + 2)
    dmx = np.zeros(nm + 2)
    dx[1:nm + 1] = x[2:nm + 2] - x[1:nm + 1]
    if pbc:
        dx[0], dx[nm + 1] = dx[nm], dx[1]
    dmx[1:nm + 1] = 0.5 * (dx[0:nm] + dx[1:nm + 1])
    if pbc:
        dmx[0], dmx[nm + 1] = dmx[nm], dmx[1]
    else:
        dmx[nm + 1] = 0.5 * (dx[nm] + dx[nm + 1])
    return Axis(x=x, dx=dx, dmx=dmx, nm=nm, L=L, pbc=pbc, uniform=uniform, gamma=gamma)


def cell_centres(ax: Axis) -> np.ndarray:
    """Interior cell-centre coordinates (x_i + x_{i+1})/2, i = 1..nm."""
    return 0.5 * (ax.x[1:ax.nm + 1] + ax.x[2:ax.nm + 2])
