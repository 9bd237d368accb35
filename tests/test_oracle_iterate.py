"""Pins of the oracle's ChannelIterate sequence (src/m_ChannelSystem.f90:80-140)
by what the mathematics of the scheme fixes, on the CPU.

The per-procedure pins elsewhere fix each stencil, solve and boundary fill;
these fix how ChannelIterate composes them: which substep applies which of
gamma, zeta and alpha (src/m_Parameters.f90:159-174,215-230), that the
explicit history is carried from substep to substep and from step to step,
that the Crank-Nicolson half-weight beta = alpha nu dt / 2 sits on both the
explicit and the implicit side, the constant-flow-rate forcing of both builds
(src/m_TScheme.f90:269-300) and the PrGradAver bookkeeping.

* Separable modes.  u = phi(y) cos(k z) with v = w = p = 0 (or w = phi(y)
  cos(k x), or a scalar c = chi(y) cos(k x) at rest) has no convection and no
  divergence, so a step only applies the explicit x/z diffusion (a Fourier
  symbol times the mode) and the y diffusion.  When phi is an eigenvector of
  the wall-normal operator, built here from the raw second-difference rows and
  the ghost-cell identity of the wall condition, every substep multiplies the
  mode's amplitude and its history by a 2 x 2 matrix of the RK coefficients:
  the oracle's field after one or several steps must be that closed form.
* Flows that depend on y only.  Convection and the projection vanish again,
  and the step is a 1-D recurrence with dense solves, written here from the
  scheme's definition.  The discrete Poiseuille profile nu L_y U = -G at the
  required bulk velocity must stay fixed for 100 steps with PrGradAver = -G
  (non-scalar build) and the moving-frame Poiseuille profile with a body
  force (ScalarFlow build); a non-steady profile must follow the recurrence.
* Self-convergence in dt.  A smooth, projected, nonlinear channel state run to
  a fixed time with dt, dt/2 and dt/4 must converge at second order towards a
  run with dt/32.
"""
import math

import numpy as np
import pytest

from helpers import oracle_project
from oracle import oracle as O
from paratc_amd import case as CS, ic

# the RK tables of the three time schemes (gamma, zeta), src/m_Parameters.f90:159-174
SCHEMES = {
    1: ((1.5,), (-0.5,)),                                        # AB2
    2: ((0.5, 1.0), (0.0, -0.5)),                                # RK2
    3: ((8 / 15, 5 / 12, 3 / 4), (0.0, -17 / 60, -5 / 12)),      # RK3
}


def halo(core):
    """Haloed array from a core (nx, ny+2, nz) with periodic x/z images."""
    return ic._halo_from_core(np.asfortranarray(core))


def wall_operator(o, rows, ny, r_bot, r_top):
    """The wall-normal second difference as an ny x ny matrix: the raw rows
    (am, ac, ap) of a cell-centred variable, with the ghost values of the wall
    conditions folded in as ghost = r * (first interior value) (homogeneous
    conditions)."""
    am, ac, ap = (o.metric(n)[1:ny + 1] for n in rows)
    A = np.diag(ac) + np.diag(am[1:], -1) + np.diag(ap[:-1], 1)
    A[0, 0] += r_bot * am[0]
    A[-1, -1] += r_top * ap[-1]
    return A


def eigenmodes(A):
    lam, vec = np.linalg.eig(A)
    assert np.abs(lam.imag).max() < 1e-9 * np.abs(lam.real).max()
    order = np.argsort(-lam.real)  # smooth modes first
    return lam.real[order], vec.real[:, order]


def d2_symbol(m, n, length):
    """Symbol of the 4th-order second difference (-1, 16, -30, 16, -1) / 12 dx^2
    on cos(2 pi m x / L)."""
    w = 2.0 * math.pi * m / n
    rd2 = (n / length) ** 2
    return rd2 * (-2.0 * math.cos(2 * w) + 32.0 * math.cos(w) - 30.0) / 12.0


def mode_amplitude(ischeme, dt, diff, mu, lam, nsteps, a0=1.0, h0=0.0):
    """Amplitude of a separable mode after nsteps: per substep the explicit
    part H = diff * mu * a, Rhs = gamma dt H + zeta dt H_old + 2 beta lam a with
    beta = alpha diff dt / 2, da = Rhs / (1 - beta lam), H_old <- H."""
    gam, zet = SCHEMES[ischeme]
    a, h = a0, h0
    for _ in range(nsteps):
        for g, z in zip(gam, zet):
            al = g + z
            beta = 0.5 * al * diff * dt
            H = diff * mu * a
            a = a + (g * dt * H + z * dt * h + 2.0 * beta * lam * a) / (1.0 - beta * lam)
            h = H
    return a


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def u_mode_state(case, o, phi, m, direction):
    """u = phi(y) cos(2 pi m k / nz) (direction 'u') or w = phi(y) cos(2 pi m i / nx)
    (direction 'w'), other velocities and p zero, walls filled."""
    nx, ny, nz = case.nx, case.ny, case.nz
    core = np.zeros((nx, ny + 2, nz))
    if direction == "u":
        core[:, 1:ny + 1, :] = phi[None, :, None] * np.cos(2 * np.pi * m * np.arange(nz) / nz)[None, None, :]
    else:
        core[:, 1:ny + 1, :] = phi[None, :, None] * np.cos(2 * np.pi * m * np.arange(nx) / nx)[:, None, None]
    st = O.State(o)
    if direction == "u":
        st.ux = halo(core)
    else:
        st.uz = halo(core)
    o.SetBC_and_UpdateHalo(st.ux, st.uy, st.uz, st.uy_ym)
    o.SetBC_and_UpdateHalo_pr(st.p)
    return st, core


@pytest.mark.parametrize("ischeme,nsteps", [(3, 1), (3, 3), (2, 2), (1, 3)])
@pytest.mark.parametrize("flow_type", [CS.FT_CH, CS.FT_HC])
@pytest.mark.parametrize("direction,mode", [("u", 0), ("u", 3), ("w", 1)])
def test_separable_velocity_mode(ischeme, nsteps, flow_type, direction, mode):
    nx, ny, nz = 16, 12, 16
    case = CS.case_channel(nx, ny, nz, 1.6, re_b=300.0, dt=4e-3)
    case.flow_type = flow_type
    if flow_type == CS.FT_HC:
        case.sections = [CS.MeshSection(1.0, ny, 1, 0, 1.6)]
        case.yly = 1.0
    case.is_ux_const = False
    case.ischeme = ischeme
    o = O.Oracle(case)
    # u_ghost = -u_1 at a no-slip wall, u_ghost = u_ny at the free-slip top
    A = wall_operator(o, ("am2c", "ac2c", "ap2c"), ny, -1.0, -1.0 if flow_type == CS.FT_CH else 1.0)
    lam, vec = eigenmodes(A)
    mz = 2  # z wavenumber of the u mode
    jm = mode
    phi = vec[:, jm] / np.abs(vec[:, jm]).max()
    st, core = u_mode_state(case, o, phi, mz if direction == "u" else 3, direction)
    mu = d2_symbol(mz, nz, case.zlz) if direction == "u" else d2_symbol(3, nx, case.xlx)
    for _ in range(nsteps):
        o.ChannelIterate(st, 1)
    a = mode_amplitude(ischeme, case.dtmax, case.xnu, mu, lam[jm], nsteps)
    expect = halo(core * a)
    got = st.ux if direction == "u" else st.uz
    other = st.uz if direction == "u" else st.ux
    inner = (slice(2, -2), slice(1, -1), slice(2, -2))
    e = rel(got[inner], expect[inner])
    assert e < 1e-13, e
    # the amplification is not trivial: diffusion moved the amplitude
    assert abs(a - 1.0) > 1e-4
    for f in (other, st.uy, st.p):
        assert np.abs(f[inner]).max() < 1e-12 * abs(a)  # round-off divergence of the mode


@pytest.mark.parametrize("opts,vals", [((-1, -1), (0.0, 0.0, 0.0, 0.0)),
                                       ((-2, -2), (0.0, 0.0, 0.0, 0.0)),
                                       ((-3, -1), (7.0, 0.0, 0.0, 0.0)),
                                       ((-2, -3), (0.0, 40.0, 0.0, 0.0))])
@pytest.mark.parametrize("ischeme,nsteps", [(3, 2), (1, 2)])
def test_separable_scalar_mode(opts, vals, ischeme, nsteps):
    """c = chi(y) cos(2 pi 2 i / nx) at rest, no gravity coupling and no
    settling: the Dirichlet / Neumann / Robin walls enter through the ghost
    factor r of SetBC_and_UpdateHalo_scalar (src/m_BC_and_Halo.f90:128-178)."""
    nx, ny, nz = 16, 12, 8
    case = CS.case_gravity_current(nx, ny, nz, gamma=1.5, re=200.0, ws=0.0, ri=0.0, dt=5e-3)
    case.scalar_bc_option = opts
    case.scalar_bc_values = vals
    case.schmidt = 0.7
    case.ischeme = ischeme
    o = O.Oracle(case)
    # dyc(1) and dyc(nyp) span the mirrored ghost centre and the first/last
    # centre: twice the wall-to-centre distance (src/m_MeshAndMetries.f90:195-224)
    yp = case.yp()
    yc = 0.5 * (yp[:-1] + yp[1:])
    dyc1 = 2.0 * (yc[0] - yp[0])
    dycn = 2.0 * (yp[-1] - yc[-1])

    def r_of(opt, F, dyc, top):
        if opt == -1:
            return -1.0
        if opt == -2:
            return 1.0
        return (2.0 - F * dyc) / (2.0 + F * dyc) if top else (2.0 + F * dyc) / (2.0 - F * dyc)
    dyc = o.metric("dyc")
    assert abs(dyc[1] - dyc1) < 1e-14 and abs(dyc[ny + 1] - dycn) < 1e-14
    rb = r_of(opts[0], vals[0], dyc[1], False)
    rt = r_of(opts[1], vals[1], dyc[ny + 1], True)
    A = wall_operator(o, ("am2cSc", "ac2cSc", "ap2cSc"), ny, rb, rt)
    lam, vec = eigenmodes(A)
    jm = 1
    chi = vec[:, jm] / np.abs(vec[:, jm]).max()
    mx = 2
    core = np.zeros((nx, ny + 2, nz))
    core[:, 1:ny + 1, :] = chi[None, :, None] * np.cos(2 * np.pi * mx * np.arange(nx) / nx)[:, None, None]
    st = O.State(o)
    st.scalar = halo(core)
    o.SetBC_and_UpdateHalo(st.ux, st.uy, st.uz, st.uy_ym)
    o.SetBC_and_UpdateHalo_pr(st.p)
    o.SetBC_and_UpdateHalo_scalar(st.scalar)
    for _ in range(nsteps):
        o.ChannelIterate(st, 1)
    a = mode_amplitude(ischeme, case.dtmax, case.sdiff, d2_symbol(mx, nx, case.xlx), lam[jm], nsteps)
    inner = (slice(2, -2), slice(1, -1), slice(2, -2))
    e = rel(st.scalar[inner], halo(core * a)[inner])
    assert e < 1e-13, e
    assert abs(a - 1.0) > 1e-4
    for f in (st.ux, st.uy, st.uz, st.p):
        assert np.abs(f[inner]).max() < 1e-14


# ------------------------------------------------------------ y-only flows
def y_only_case(scalar, ucrf=0.0, ischeme=3):
    nx, ny, nz = 8, 16, 8
    if scalar:
        case = CS.case_gravity_current(nx, ny, nz, gamma=1.7, re=400.0, ws=0.0, ri=0.0, dt=3e-3)
        case.scalar_bc_option = (-1, -1)
        case.scalar_bc_values = (0.0, 0.0, 0.0, 0.0)
        case.is_use_crf, case.ucrf_in = ucrf != 0.0, ucrf
    else:
        case = CS.case_channel(nx, ny, nz, 1.7, re_b=400.0, dt=3e-3)
    case.is_ux_const = True
    case.ischeme = ischeme
    return case


def y_only_step(case, o, U, h, G, ubulk):
    """One step of ChannelIterate on u = U(y), v = w = p = 0, written from the
    scheme: no convection, no divergence; explicit H = gravity(1); the
    Crank-Nicolson y-diffusion with the wall ghosts u_0 = 2 u_w - u_1; the
    constant-flow-rate forcing of the build.  Returns U, H_old, PrGradAver."""
    ny = case.ny
    nu, dt = case.xnu, case.dtmax
    A = wall_operator(o, ("am2c", "ac2c", "ap2c"), ny, -1.0, -1.0)
    am, ap = o.metric("am2c")[1:ny + 1], o.metric("ap2c")[1:ny + 1]
    uw = -case.ucrf  # both walls at rest in the lab frame
    wall = np.zeros(ny)
    wall[0] += 2.0 * uw * am[0]
    wall[-1] += 2.0 * uw * ap[-1]
    dyp = o.metric("dyp")[1:ny + 1]
    gam, zet = SCHEMES[case.ischeme]
    prgrad = 0.0
    for g, z in zip(gam, zet):
        al = g + z
        beta = 0.5 * al * nu * dt
        H = np.full(ny, G)
        LU = A @ U + wall
        rhs = g * dt * H + z * dt * h + 2.0 * beta * LU
        if case.scalar:
            dpm = ubulk - case.ucrf - (U * dyp).sum() / case.yly
            rhs = rhs + al * dpm
            prgrad += dpm * al / dt
        else:
            dpm = nu * (LU * dyp).sum() / case.yly
            rhs = rhs - al * dt * dpm
            prgrad += dpm * al
        U = U + np.linalg.solve(np.eye(ny) - beta * A, rhs)
        h = H
    return U, h, prgrad


def run_y_only(case, U0, G, nsteps):
    o = O.Oracle(case)
    nx, ny, nz = case.nx, case.ny, case.nz
    core = np.zeros((nx, ny + 2, nz))
    core[:, 1:ny + 1, :] = U0[None, :, None]
    st = O.State(o)
    st.ux = halo(core)
    o.SetBC_and_UpdateHalo(st.ux, st.uy, st.uz, st.uy_ym)
    if case.scalar:
        o.SetBC_and_UpdateHalo_scalar(st.scalar)
    prg = []
    for _ in range(nsteps):
        o.ChannelIterate(st, 1)
        prg.append(o.coe()[8])
    return o, st, prg


def poiseuille(case, o, G):
    """Discrete Poiseuille profile: nu (L_y U + wall terms) = -G."""
    ny = case.ny
    A = wall_operator(o, ("am2c", "ac2c", "ap2c"), ny, -1.0, -1.0)
    am, ap = o.metric("am2c")[1:ny + 1], o.metric("ap2c")[1:ny + 1]
    wall = np.zeros(ny)
    wall[0] += 2.0 * (-case.ucrf) * am[0]
    wall[-1] += 2.0 * (-case.ucrf) * ap[-1]
    return np.linalg.solve(A, -G / case.xnu - wall)


def test_discrete_poiseuille_is_steady_nonscalar_build():
    """IsUxConst, non-ScalarFlow build (src/m_TScheme.f90:287-299): the
    profile at the required bulk velocity stays put for 100 steps and the
    mean pressure gradient is PrGradAver = -G with nu L_y U = -G."""
    case = y_only_case(False)
    o = O.Oracle(case)
    dyp = o.metric("dyp")[1:case.ny + 1]
    U1 = poiseuille(case, o, 1.0)
    bulk1 = (U1 * dyp).sum() / case.yly
    G = case.ubulk / bulk1
    U = U1 * G
    o2, st, prg = run_y_only(case, U, 0.0, 100)
    got = st.ux[2, 1:-1, 2]
    assert rel(got, U) < 1e-12
    assert np.abs(st.ux[2:-2, 1:-1, 2:-2] - got[None, :, None]).max() < 1e-13
    assert abs(prg[-1] - (-G)) < 1e-11 * G
    assert abs(prg[0] - (-G)) < 1e-11 * G
    for f in (st.uy, st.uz, st.p):
        assert np.abs(f[2:-2, 1:-1, 2:-2]).max() < 1e-13


@pytest.mark.parametrize("ucrf", [0.0, 0.35])
def test_discrete_poiseuille_is_steady_scalar_build(ucrf):
    """IsUxConst, ScalarFlow build (src/m_TScheme.f90:269-285), driven by a
    body force gravity(1) = G in the frame moving at uCRF: the profile with
    bulk velocity ubulk - uCRF stays put and PrGradAver stays at zero."""
    case = y_only_case(True, ucrf)
    G = 2.0e-3
    case.gravity = (G, 0.0, 0.0)
    o = O.Oracle(case)
    dyp = o.metric("dyp")[1:case.ny + 1]
    U = poiseuille(case, o, G)
    case.ubulk = (U * dyp).sum() / case.yly + case.ucrf
    o2, st, prg = run_y_only(case, U, G, 100)
    got = st.ux[2, 1:-1, 2]
    assert rel(got, U) < 1e-12
    assert max(abs(v) for v in prg) < 1e-10
    for f in (st.uy, st.uz, st.p, st.scalar):
        assert np.abs(f[2:-2, 1:-1, 2:-2]).max() < 1e-13


@pytest.mark.parametrize("scalar,ucrf,ischeme", [(False, 0.0, 3), (False, 0.0, 1), (False, 0.0, 2),
                                                 (True, 0.0, 3), (True, 0.25, 3), (True, 0.25, 2)])
def test_y_only_transient_follows_recurrence(scalar, ucrf, ischeme):
    """A profile that is not steady: U(y) = a Poiseuille-like bump plus a
    wall-to-wall tilt, pulled towards the required bulk velocity by the
    constant-flow-rate forcing.  Five steps against the recurrence; the
    history, the forcing and PrGradAver (reset at the first substep) all
    enter."""
    case = y_only_case(scalar, ucrf, ischeme)
    G = 1.5e-3 if scalar else 0.0
    case.gravity = (G, 0.0, 0.0)
    o = O.Oracle(case)
    yp = case.yp()
    yc = 0.5 * (yp[:-1] + yp[1:])
    U0 = 1.2 * (1.0 - (yc - 1.0) ** 2) + 0.3 * (yc - 1.0) + 0.1 * np.cos(3.0 * yc)
    case.ubulk = 0.9
    o2, st, prg = run_y_only(case, U0, G, 5)
    U, h = U0.copy(), np.zeros(case.ny)
    for s in range(5):
        U, h, pg = y_only_step(case, o2, U, h, G, case.ubulk)
        assert abs(prg[s] - pg) < 1e-11 * max(1.0, abs(pg)), (s, prg[s], pg)
    assert rel(st.ux[2, 1:-1, 2], U) < 1e-13
    assert rel(U, U0) > 1e-4  # it did move


# ------------------------------------------------------- self-convergence
def smooth_state(case, o, amp=0.15):
    """A smooth divergence-free channel state: the laminar profile plus a few
    Fourier modes with wall-vanishing y shapes, projected once (clcPrSrc,
    clcPPE, FluidVelUpdate with alpha dt = 1)."""
    nx, ny, nz = case.nx, case.ny, case.nz
    yp = case.yp()
    yc = 0.5 * (yp[:-1] + yp[1:])
    x = (np.arange(nx) + 0.5) * case.xlx / nx
    z = (np.arange(nz) + 0.5) * case.zlz / nz
    X, Y, Z = np.meshgrid(x, yc, z, indexing="ij")
    bump = (1.0 - (Y - 1.0) ** 2)
    u = 1.5 * bump + amp * bump * np.sin(2 * np.pi * X / case.xlx) * np.cos(2 * np.pi * Z / case.zlz)
    w = amp * bump * np.cos(2 * np.pi * 2 * X / case.xlx) * np.sin(2 * np.pi * Z / case.zlz)
    Xv, Yv, Zv = np.meshgrid(x, yp[:-1], z, indexing="ij")
    v = amp * np.sin(np.pi * Yv / 2.0) ** 2 * np.sin(2 * np.pi * Xv / case.xlx) * np.sin(2 * np.pi * 2 * Zv / case.zlz)
    st = O.State(o)
    for name, core_in in (("ux", u), ("uy", v), ("uz", w)):
        core = np.zeros((nx, ny + 2, nz))
        core[:, 1:ny + 1, :] = core_in
        setattr(st, name, halo(core))
    st.uy[:, 1, :] = 0.0
    o.SetBC_and_UpdateHalo(st.ux, st.uy, st.uz, st.uy_ym)
    assert oracle_project(o, st) < 1e-12
    return st


def run_to(case, T, dt):
    case.dtmax = dt
    o = O.Oracle(case)
    st = smooth_state(case, o)
    n = int(round(T / dt))
    for _ in range(n):
        o.ChannelIterate(st, 1)
    return np.concatenate([a[2:-2, 1:-1, 2:-2].ravel() for a in (st.ux, st.uy, st.uz)])


@pytest.mark.parametrize("ischeme", [3, 2])
def test_self_convergence_in_dt(ischeme):
    """Second order in time for the velocities (semi-implicit RK with
    Crank-Nicolson y-diffusion and one projection per substep): the errors at
    dt, dt/2 and dt/4 against dt/32 fall by >= 2^1.8 per halving.  A zeta on
    the wrong substep, a stale history or beta = alpha nu dt would leave a
    first-order or inconsistent scheme."""
    case = CS.case_channel(16, 12, 16, 1.5, re_b=150.0)
    case.ischeme = ischeme
    T, dt0 = 0.24, 0.04
    ref = run_to(case, T, dt0 / 32)
    errs = [float(np.abs(run_to(case, T, dt0 / 2 ** q) - ref).max()) for q in range(3)]
    orders = [math.log2(errs[q] / errs[q + 1]) for q in range(2)]
    print("errors", errs, "orders", orders)
    assert min(orders) >= 1.8, (errs, orders)
