"""Seeded initial fields in the reference's haloed layout.

The reference draws its initial perturbation from a clock-seeded generator
(InitVelocity, src/m_FlowCase.f90:49-107), so a run cannot be reproduced.
These builders give the same kind of field from a fixed seed:

* ``channel_fields``: u = 1.5 (1 - (y - 1)^2) at the u points (cell centres in
  y) plus uniform noise in [-amp, amp] on u, v and w; v = 0 on the walls.
* ``gravity_current_fields``: c = 0.5 (1 - tanh((x - x0) / delta)) plus noise,
  velocities of noise only.

Arrays are Fortran-ordered (nx+4, ny+2, nz+4); x and z halos hold the periodic
images, y ghost rows are zero until the boundary routines fill them.
"""
from __future__ import annotations

import numpy as np

from . import rng


def _halo_from_core(core: np.ndarray) -> np.ndarray:
    nx, nyh, nz = core.shape
    ii = (np.arange(-1, nx + 3) - 1) % nx
    kk = (np.arange(-1, nz + 3) - 1) % nz
    return np.asfortranarray(core[ii][:, :, kk])


def _noise(seed: int, stream: int, shape, amp: float) -> np.ndarray:
    n = int(np.prod(shape))
    return amp * (2.0 * rng.uniform(seed, n, stream) - 1.0).reshape(shape, order="F")


def channel_fields(case, seed: int = 0, amp: float = 0.1):
    nx, ny, nz = case.nx, case.ny, case.nz
    yp = case.yp()
    yc = 0.5 * (yp[:-1] + yp[1:])
    h = 0.5 * case.yly
    base = 1.5 * (1.0 - ((yc - h) / h) ** 2) * case.ubulk
    ux = np.zeros((nx, ny + 2, nz), order="F")
    uy = np.zeros_like(ux)
    uz = np.zeros_like(ux)
    ux[:, 1:ny + 1, :] = base[None, :, None] + _noise(seed, 0, (nx, ny, nz), amp)
    uy[:, 1:ny + 1, :] = _noise(seed, 1, (nx, ny, nz), amp)
    uy[:, 1, :] = 0.0  # bottom wall face
    uz[:, 1:ny + 1, :] = _noise(seed, 2, (nx, ny, nz), amp)
    return _halo_from_core(ux), _halo_from_core(uy), _halo_from_core(uz)


def gravity_current_fields(case, seed: int = 2, amp: float = 1e-3, x0: float = 2.0, delta: float = 0.05):
    nx, ny, nz = case.nx, case.ny, case.nz
    dx = case.xlx / nx
    xc = (np.arange(nx) + 0.5) * dx
    c0 = 0.5 * (1.0 - np.tanh((xc - x0) / delta))
    s = np.zeros((nx, ny + 2, nz), order="F")
    s[:, 1:ny + 1, :] = c0[:, None, None] + _noise(seed, 3, (nx, ny, nz), amp)
    ux = np.zeros_like(s)
    uy = np.zeros_like(s)
    uz = np.zeros_like(s)
    ux[:, 1:ny + 1, :] = _noise(seed, 0, (nx, ny, nz), amp)
    uy[:, 1:ny + 1, :] = _noise(seed, 1, (nx, ny, nz), amp)
    uy[:, 1, :] = 0.0
    uz[:, 1:ny + 1, :] = _noise(seed, 2, (nx, ny, nz), amp)
    return _halo_from_core(ux), _halo_from_core(uy), _halo_from_core(uz), _halo_from_core(s)


def poisson_rhs(case, seed: int = 1, stream: int = 0) -> np.ndarray:
    """i.i.d. N(0,1) per cell minus its dyp-weighted mean (compatible with the
    Neumann walls: sum_j dyp_j <f>_j = 0)."""
    nx, ny, nz = case.nx, case.ny, case.nz
    f = rng.normal(seed, nx * ny * nz, stream).reshape((nx, ny, nz), order="F")
    yp = case.yp()
    dyp = yp[1:] - yp[:-1]
    m = (f.sum(axis=(0, 2)) * dyp).sum() / (dyp.sum() * nx * nz)
    return np.asfortranarray(f - m)
