"""Pins of the oracle's clcStat accumulation (src/m_FlowCase.f90:283-413).

Constant fields fix which slot holds which moment (and where uCRF enters);
single Fourier modes and linear profiles fix the derivative-based slots
through the staggered symbols; the last plane j = ny+1 carries only slots
27, 33 and 35 (1-based).
"""
import math

import numpy as np

from oracle import oracle as O
from paratc_amd import case as CS
from paratc_amd import ensemble as E
from test_oracle_rhs import Grid, si_lam


def stats(case, ux, uy, uz, p, sc=None):
    o = O.Oracle(case)
    ns = 40 if case.scalar else 35
    out = np.zeros((ns, case.ny + 1), order="F")
    o.clcStat(ux, uy, uz, p, sc if sc is not None else np.zeros_like(ux), out)
    return o, out


def test_constant_fields_fill_the_moment_slots():
    case = E.case_channel_scalar(16, 12, 10, re_b=3000.0, ws=0.02)
    case.is_use_crf, case.ucrf_in = True, 0.3
    o = O.Oracle(case)
    g = Grid(case, o)
    U, V, W, P, S = 0.7, -0.2, 0.4, 1.3, 0.9
    _, st = stats(case, g.field() * U, g.field() * V, g.field() * W, g.field() * P, g.field() * S)
    u = U + case.ucrf
    expect = [u, V, W, P, u * u, V * V, W * W, P * P, u * V, V * W, u * W, u * P, V * P, W * P,
              u * u * V, V ** 3, W * W * V, u * V * V, u ** 3, W ** 3, u ** 4, V ** 4, W ** 4,
              0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, S, S * S, S * u, S * V, S * W]
    inner = st[:, :-1]
    for q, e in enumerate(expect):
        assert np.allclose(inner[q], e, rtol=1e-12, atol=1e-11), q
    assert np.allclose(st[:, -1], 0.0, atol=1e-11)


def test_derivative_slots_by_symbols():
    """u = A sin(kz z): <u^2> = A^2/2, <u d2u/dz2> = A^2 s_zz / 2 and
    <vor_y^2> = <(du/dz)^2> = A^2 lam_z^2 / 2; with p = P and v = b y:
    <p dv/dy> = P b."""
    case = CS.case_channel(16, 12, 20, 1.6)
    o = O.Oracle(case)
    g = Grid(case, o)
    A, P, b = 0.8, 0.6, -0.45
    kz = 3 * 2 * math.pi / case.zlz
    _, lam = si_lam(kz, g.dz)
    w = kz * g.dz
    szz = (-math.cos(2 * w) + 16 * math.cos(w) - 15) / 6 / g.dz ** 2
    ux = g.field(fz=A * np.sin(kz * g.zc))
    ux[:, 0, :] = ux[:, 1, :]  # constant in y, ghosts included
    ux[:, -1, :] = ux[:, -2, :]
    uy = g.field(fy=b * g.yf)
    _, st = stats(case, ux, uy, g.zero(), g.field() * P)
    inner = st[:, :-1]
    assert np.allclose(inner[4], A * A / 2, rtol=1e-12)
    assert np.allclose(inner[27], A * A * szz / 2, rtol=1e-11)
    assert np.allclose(inner[33], A * A * lam * lam / 2, rtol=1e-11)
    assert np.allclose(inner[24], P * b, rtol=1e-12)
    assert np.allclose(inner[23], 0.0, atol=1e-12) and np.allclose(inner[25], 0.0, atol=1e-12)


def test_top_plane_wall_shear():
    """Plane ny+1 (:393-413): <(du/dy)^2> from the top ghost row."""
    case = CS.case_channel(8, 12, 8, 1.6)
    o = O.Oracle(case)
    g = Grid(case, o)
    ux = g.zero()
    ux[:, case.ny, :] = 0.25  # last interior row
    ux[:, case.ny + 1, :] = -0.25  # its no-slip image
    _, st = stats(case, ux, g.zero(), g.zero(), g.zero())
    rdyc = o.metric("rdyc")[case.ny + 1]
    assert math.isclose(st[34, -1], (0.5 * rdyc) ** 2, rel_tol=1e-12)
    assert st[32, -1] == 0.0 and st[26, -1] == 0.0
