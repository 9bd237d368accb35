"""Seeded initial fields in the reference's haloed layout.

The reference draws its initial perturbation from a clock-seeded generator
(InitVelocity, src/m_FlowCase.f90:49-107), so a run cannot be reproduced.
These builders give the same kind of field from a fixed seed:

* ``channel_fields``: u = 1.5 (1 - (y - 1)^2) at the u points (cell centres in
  y) plus uniform noise in [-amp, amp] on u, v and w; v = 0 on the walls.
* ``gravity_current_fields``: c = 0.5 (1 - tanh((x - x0) / delta)) plus noise,
  velocities of noise only.
* ``init_velocity``: InitVelocity itself, its Deviation factor in [0.9, 1.1]
  drawn from the seeded generator instead of the clock.

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


def init_velocity(case, seed: int = 0):
    """InitVelocity (src/m_FlowCase.f90:49-107): the laminar-like profile
    3 ubulk (ybar - ybar^2 / 2) plus streaks ubulk ybar exp(-4.5 ybar^2) times
    cos(wz z+) in u and sin(wx x+) in w, both scaled by the point's Deviation
    0.2 U + 0.9 (U uniform in [0, 1), here seeded), wavelengths 500 and 200
    wall units rounded to fit the box (Re_tau from the 0.1538 Re_m^0.887741
    correlation).  u is then scaled to the bulk velocity (CalcUxAver, the
    dyp-weighted mean) and shifted by -uCRF, and w loses its x-z plane means;
    v = 0."""
    nx, ny, nz = case.nx, case.ny, case.nz
    yp = case.yp()
    yc = 0.5 * (yp[:-1] + yp[1:])
    dyp = yp[1:] - yp[:-1]
    height = 0.5 * case.yly if case.flow_type == 1 else case.yly
    rem = case.ubulk * height / case.xnu
    utau = 0.1538 * rem ** 0.887741 * case.xnu / height
    twopi = 2.0 * np.pi
    xlx_plus, zlz_plus = case.xlx * utau / case.xnu, case.zlz * utau / case.xnu
    wx = (np.floor(xlx_plus * (twopi / 500.0) / twopi) + 1.0) * twopi / xlx_plus
    wz = (np.floor(zlz_plus * (twopi / 200.0) / twopi) + 1.0) * twopi / zlz_plus
    dx, dz = case.xlx / nx, case.zlz / nz
    xp = np.arange(nx) * dx + 0.5 * dx
    zp = np.arange(nz) * dz + 0.5 * dz
    dev = 0.2 * rng.uniform(seed, nx * ny * nz, 0).reshape((nx, ny, nz), order="F") + 0.9
    ybar = (height - np.abs(height - yc)) / height
    amp = case.ubulk * ybar * np.exp(-4.5 * ybar * ybar)
    ux_c = amp[None, :, None] * np.cos(wz * utau * zp / case.xnu)[None, None, :] * dev
    uz_c = amp[None, :, None] * np.sin(wx * utau * xp / case.xnu)[:, None, None] * dev
    ux_c = ux_c + (3.0 * case.ubulk * (ybar - 0.5 * ybar * ybar))[None, :, None]
    uzmean = uz_c.mean(axis=(0, 2))
    ux_aver = float((ux_c * dyp[None, :, None]).sum()) / (nx * nz) / case.yly
    ux_c = ux_c * (case.ubulk / ux_aver) - case.ucrf
    uz_c = uz_c - uzmean[None, :, None]
    ux = np.zeros((nx, ny + 2, nz), order="F")
    uy = np.zeros_like(ux)
    uz = np.zeros_like(ux)
    ux[:, 1:ny + 1, :] = ux_c
    uz[:, 1:ny + 1, :] = uz_c
    return _halo_from_core(ux), _halo_from_core(uy), _halo_from_core(uz)
