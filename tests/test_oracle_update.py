"""Pins of the oracle's projection updates, PrSrc boundary routine and monitors.

FluidVelUpdate / PressureUpdate (src/m_TScheme.f90:82-107,512-532) against
Fourier symbols and polynomial exactness; SetBC_and_UpdateHaloForPrSrc
(src/m_BC_and_Halo.f90:69-103) and Update_uy_ym (src/m_FlowCase.f90:112-122)
against their defining assignments checked with numpy; the monitors of
src/m_Tools.f90:24-110,256-301 against numpy on fields whose extremes are
known.
"""
import math

import numpy as np

from helpers import rel_err
from oracle import oracle as O
from paratc_amd import case as CS
from test_oracle_rhs import Grid, si_lam


def channel(nx=24, ny=16, nz=20, gamma=1.7, uniform=False):
    case = CS.case_channel(nx, ny, nz, gamma, dt=1e-2)
    if uniform:
        case.sections = [CS.MeshSection(2.0, ny, 0, 0, 0.0)]
    return case


def test_fluid_vel_update_is_minus_alpha_grad_phi():
    """u -= alpha G phi: a cosine in x and z gives -lam sin at the faces, and a
    linear profile in y gives its slope exactly on the stretched mesh."""
    case = channel()
    o = O.Oracle(case)
    g = Grid(case, o)
    o.PMcoeUpdate(2, 1e-2)
    alpha = o.coe()[2]
    kx, kz, b = 2 * (2 * math.pi / case.xlx), 3 * (2 * math.pi / case.zlz), -0.7
    _, lx = si_lam(kx, g.dx)
    _, lz = si_lam(kz, g.dz)
    phi = g.field(fx=np.cos(kx * g.xc)) + g.field(fz=np.cos(kz * g.zc)) + g.field(fy=b * g.yc)
    ux, uy, uz = g.zero(), g.zero(), g.zero()
    o.FluidVelUpdate(phi, ux, uy, uz)
    ex = -alpha * -lx * np.sin(kx * g.xf[2:-2])
    ez = -alpha * -lz * np.sin(kz * g.zf[2:-2])
    assert rel_err(ux[2:-2, 1:-1, 2:-2], np.broadcast_to(ex[:, None, None], (case.nx, case.ny, case.nz))) < 1e-12
    assert rel_err(uz[2:-2, 1:-1, 2:-2], np.broadcast_to(ez[None, None, :], (case.nx, case.ny, case.nz))) < 1e-12
    # rows j >= 2: (phi_j - phi_{j-1}) / (yc_j - yc_{j-1}); row 1 uses dyc(1) = dyp(1)
    assert rel_err(uy[2:-2, 2:-1, 2:-2], np.full((case.nx, case.ny - 1, case.nz), -alpha * b)) < 1e-12


def test_pressure_update_adds_phi_minus_beta_lap_y_phi():
    """p += (1 - beta L_y) phi: for phi constant in y the Neumann operator
    vanishes in every row; for a quadratic on a uniform mesh it is phi''
    in the interior rows."""
    case = channel(uniform=True)
    o = O.Oracle(case)
    g = Grid(case, o)
    o.PMcoeUpdate(3, 1e-2)
    beta = o.coe()[4]
    kx = 2 * math.pi / case.xlx
    phi = g.field(fx=np.cos(kx * g.xc))
    p = g.field() * 0.3
    o.PressureUpdate(p, phi)
    assert rel_err(p[2:-2, 1:-1, 2:-2], 0.3 + phi[2:-2, 1:-1, 2:-2]) < 1e-14
    c2 = 0.45
    phi = g.field(fy=1.0 + 0.2 * g.yc + c2 * g.yc ** 2)
    p = g.zero()
    o.PressureUpdate(p, phi)
    inner_rows = p[2:-2, 2:-2, 2:-2]
    assert rel_err(inner_rows, phi[2:-2, 2:-2, 2:-2] - beta * 2 * c2) < 1e-12


def test_set_bc_for_prsrc():
    """v on the wall faces (uy_ym below, the top uyBcValue above) and the x
    halo (1, 2) of u and the z halo (1, 2) of w filled periodically."""
    case = channel()
    case.uybc = (0, 0, 0, 0.125, 0, 0)
    o = O.Oracle(case)
    rng = np.random.default_rng(21)
    ux, uy, uz = (np.asfortranarray(rng.standard_normal((case.nx + 4, case.ny + 2, case.nz + 4))) for _ in range(3))
    ux0, uy0, uz0 = ux.copy(), uy.copy(), uz.copy()
    uy_ym = np.asfortranarray(rng.standard_normal((case.nx, case.nz)))
    o.SetBC_and_UpdateHaloForPrSrc(ux, uy, uz, uy_ym)
    ny = case.ny
    assert np.array_equal(uy[2:-2, 1, 2:-2], uy_ym)
    assert np.all(uy[2:-2, ny + 1, 2:-2] == 0.125)
    # x halo of u: one cell below, two above; w: z halo one below, two above
    assert np.array_equal(ux[1, :, 2:-2], ux0[-3, :, 2:-2])
    assert np.array_equal(ux[-2:, :, 2:-2], ux0[2:4, :, 2:-2])
    assert np.array_equal(ux[0], ux0[0])
    assert np.array_equal(uz[2:-2, :, 1], uz0[2:-2, :, -3])
    assert np.array_equal(uz[2:-2, :, -2:], uz0[2:-2, :, 2:4])
    assert np.array_equal(uz[:, :, 0], uz0[:, :, 0])


def test_update_uy_ym_resets_the_transpiration_plane():
    case = channel()
    o = O.Oracle(case)
    rng = np.random.default_rng(22)
    uy_ym = np.asfortranarray(rng.standard_normal((case.nx, case.nz)))
    old = uy_ym.copy()
    duy = np.zeros_like(uy_ym)
    o.Update_uy_ym(uy_ym, duy)
    assert not uy_ym.any() and np.array_equal(duy, -old)


def test_cfl_of_a_uniform_stream():
    """CFL = max over cells of the face-averaged |u_i| / spacing (src/m_Tools.f90:24-52):
    for constant (U, V, W) it is |U|/dx + |V|/min(dyp) + |W|/dz."""
    case = channel()
    o = O.Oracle(case)
    g = Grid(case, o)
    U, V, W = 1.3, -0.4, 0.7
    cfl = o.CalcMaxCFL(g.field() * U, g.field() * V, g.field() * W)
    dyp = np.diff(case.yp())
    ref = abs(U) / g.dx + abs(V) / dyp.min() + abs(W) / g.dz
    assert abs(cfl - ref) < 1e-13 * ref


def test_divergence_monitor_and_vmax():
    """CheckDivergence (src/m_Tools.f90:54-81) for u = sin(kx x), v = b y:
    max |lam cos(kx xc) + b|; CalcVmax = max |u_i| over the interior only."""
    case = channel()
    o = O.Oracle(case)
    g = Grid(case, o)
    kx, b = 3 * (2 * math.pi / case.xlx), 0.35
    _, lam = si_lam(kx, g.dx)
    ux = g.field(fx=np.sin(kx * g.xf))
    uy = g.field(fy=b * g.yf)
    d = o.CheckDivergence(ux, uy, g.zero())
    assert abs(d - np.abs(lam * np.cos(kx * g.xc[2:-2]) + b).max()) < 1e-12 * d
    rng = np.random.default_rng(23)
    f = [np.asfortranarray(rng.standard_normal(g.shape)) for _ in range(3)]
    f[1][5, 0, 5] = 50.0  # ghost rows and halos are not part of the maximum
    f[2][0, 5, 5] = -60.0
    v = o.CalcVmax(*f)
    ref = [np.abs(a[2:-2, 1:-1, 2:-2]).max() for a in f]
    assert np.array_equal(v, ref)


def test_mean_values_are_dyp_weighted():
    """ClcMeanValue / CalcUxAver (src/m_Tools.f90:256-301): dyp-weighted volume
    means of u, c and c^2; a field constant in x/z with a known profile."""
    case = channel()
    o = O.Oracle(case)
    g = Grid(case, o)
    yp = case.yp()
    # u = y at cell centres: the dyp-weighted mean of the midpoints of each
    # cell is the exact integral of y over [0, 2] divided by 2 = 1
    ux = g.field(fy=g.yc)
    s = g.field(fy=np.full(case.ny + 2, 2.0))
    res = o.ClcMeanValue(ux, s)
    # serial sums over nx*ny*nz terms: round-off up to ~n eps
    assert np.allclose(res, [1.0, 2.0, 4.0], rtol=1e-12, atol=0)
    assert abs(o.CalcUxAver(ux) - 1.0) < 1e-12
    assert abs(yp[-1] - case.yly) < 1e-15
