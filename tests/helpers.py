"""Shared helpers of the test suite (inputs, comparisons)."""
from __future__ import annotations

import numpy as np

from paratc_amd import rng


def rel_err(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    scale = max(float(np.abs(b).max()), 1e-300)
    return float(np.abs(a - b).max()) / scale


def inner(h):
    """Interior (x 1..nx, y 1..ny, z 1..nz) of a haloed (nx+4, ny+2, nz+4) array."""
    return h[2:-2, 1:-1, 2:-2]


def inner_y(h):
    """Interior in x and z with the y ghost rows (y 0..ny+1)."""
    return h[2:-2, :, 2:-2]


def random_interior(shape, seed, stream=0):
    n = int(np.prod(shape))
    return rng.normal(seed, n, stream).reshape(shape, order="F")


def compatible_rhs(orc_or_dyp, f):
    """Remove the dyp-weighted mean so that sum_j dyp_j <f>_j = 0 (Neumann compatibility,
    src/m_MeshAndMetries.f90:239-247)."""
    dyp = orc_or_dyp
    ny = f.shape[1]
    w = dyp[1:ny + 1]
    m = (f * w[None, :, None]).sum() / (w.sum() * f.shape[0] * f.shape[2])
    return np.asfortranarray(f - m)


def random_halo(shape_interior, seed, amp=1.0, stream=0):
    """A haloed array whose interior and y-ghost rows are random; x/z halos are the
    periodic images (as after a full myupdate_halo)."""
    nx, ny, nz = shape_interior
    core = amp * rng.uniform(seed, nx * (ny + 2) * nz, stream).reshape((nx, ny + 2, nz), order="F")
    core = 2.0 * core - amp
    h = np.zeros((nx + 4, ny + 2, nz + 4), order="F")
    ii = (np.arange(-1, nx + 3) - 1) % nx
    kk = (np.arange(-1, nz + 3) - 1) % nz
    h[:, :, :] = core[ii][:, :, kk]
    return h


def apply_laplacian(o, case, p):
    """The discrete Poisson operator clcPPE inverts (src/m_Poisson.f90:383-430)
    applied to an interior field p: the modified-wavenumber symbols of the 4th-
    order x/z second differences as their 7-point stencils (1, -54, 783, -1460,
    783, -54, 1) / 576 dx^2 with periodic wrap, and am2ph / ac2ph / ap2ph in y
    with the Neumann ghosts p_0 = p_1, p_{ny+1} = p_ny.  o: the oracle (its
    metrics)."""
    ny = case.ny
    am, ac, ap = (o.metric(n)[1:ny + 1] for n in ("am2ph", "ac2ph", "ap2ph"))
    rdx2, rdz2 = (case.nx / case.xlx) ** 2, (case.nz / case.zlz) ** 2
    st = np.array([1, -54, 783, -1460, 783, -54, 1]) / 576.0
    lp = np.zeros_like(p)
    for s in range(-3, 4):
        lp += st[s + 3] * (rdx2 * np.roll(p, -s, axis=0) + rdz2 * np.roll(p, -s, axis=2))
    pm = np.concatenate([p[:, :1], p, p[:, -1:]], axis=1)
    lp += am[None, :, None] * pm[:, :-2] + ac[None, :, None] * pm[:, 1:-1] + ap[None, :, None] * pm[:, 2:]
    return lp


def oracle_project(o, st):
    """The oracle's form of Solver.project_initial: one projection of the
    velocities of `st` with alpha dt = 1 (clcPrSrc, clcPPE Choice-1,
    FluidVelUpdate and the boundary fills), coefficients restored."""
    save = o.coe()
    c = save.copy()
    c[2] = 1.0
    o.set_coe(c)
    o.SetBC_and_UpdateHaloForPrSrc(st.ux, st.uy, st.uz, st.uy_ym)
    f = o.interior()
    o.clcPrSrc(st.ux, st.uy, st.uz, f)
    phi = o.halo()
    o.clcPPE(f, phi, 1)
    o.SetBC_and_UpdateHalo_pr(phi)
    o.FluidVelUpdate(phi, st.ux, st.uy, st.uz)
    o.SetBC_and_UpdateHalo(st.ux, st.uy, st.uz, st.uy_ym)
    o.set_coe(save)
    return o.CheckDivergence(st.ux, st.uy, st.uz)
