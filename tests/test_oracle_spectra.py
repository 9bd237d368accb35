"""Pins of the oracle's energy spectra (clcStat, src/m_FlowCase.f90:430-805,
helpers :954-1375).

- Parseval: every energy or co-spectrum of a row sums to the x-z mean of the
  product it decomposes, in both directions (catches a dropped factor 2 or a
  dropped edge wavenumber).
- Single Fourier modes fix the linear-coherence spectra in closed form: a
  row A_j cos(theta n + phi_j) against the reference row gives
  A_j A_r cos(phi_j - phi_r) / 4 (real part) and A_j A_r sin(phi_j - phi_r) / 4
  (imaginary part) at the mode, which fixes the sign convention and the
  mirrored reference row of a channel's upper half.
- numpy's complex FFT gives every 1-D quantity and, rebuilt into FFTW's
  half-complex layout along x and then z, every 2-D one.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from paratc_amd import case as CS

IC = (-1.0 / 16.0, 9.0 / 16.0, 9.0 / 16.0, -1.0 / 16.0)


def haloed(core):
    """(nx, ny+2, nz) values for y rows 0..ny+1 -> the haloed (nx+4, ny+2, nz+4) array."""
    return np.asfortranarray(np.pad(core, ((2, 2), (0, 0), (2, 2)), mode="wrap"))


def random_fields(case, seed, scale=1.0):
    r = np.random.default_rng(seed)
    n = 5 if case.scalar else 4
    return [haloed(scale * r.standard_normal((case.nx, case.ny + 2, case.nz))) for _ in range(n)]


def interior(a):
    return a[2:-2, 1:-1, 2:-2]


def ux_cell(ux, ucrf):
    u = ux[:, 1:-1, 2:-2]  # x -1..nx+2 (haloed), rows 1..ny, z 1..nz
    nx = u.shape[0] - 4
    return sum(c * u[1 + s:1 + s + nx] for s, c in enumerate(IC)) + ucrf


def uy_cell(uy):
    return 0.5 * (uy[2:-2, 1:-1, 2:-2] + uy[2:-2, 2:, 2:-2])


def hc(a, axis):
    """FFTW R2HC layout along axis from numpy's FFT."""
    n = a.shape[axis]
    F = np.fft.fft(a, axis=axis)
    out = np.empty(a.shape)
    idx = [slice(None)] * a.ndim
    for p in range(n):
        idx[axis] = p
        src = [slice(None)] * a.ndim
        src[axis] = p if p <= n // 2 else n - p
        out[tuple(idx)] = F[tuple(src)].real if p <= n // 2 else F[tuple(src)].imag
    return out


def channel(nx=16, ny=12, nz=10, scalar=False, ucrf=0.0, half=False):
    case = CS.case_channel(nx, ny, nz, 1.6)
    case.scalar = scalar
    if half:
        case.flow_type = CS.FT_HC
    if ucrf:
        case.is_use_crf, case.ucrf_in = True, ucrf
    return case


@pytest.mark.parametrize("scalar", [False, True])
def test_parseval_sums(scalar):
    case = channel(16, 12, 10, scalar=scalar, ucrf=0.3)
    o = O.Oracle(case)
    f = random_fields(case, 7)
    ux, uy, uz, p = f[:4]
    sc = f[4] if scalar else None
    sx, sz = o.spectra_1d(ux, uy, uz, p, sc, jlcs=3)
    uc = case.ucrf
    prods = {1: (interior(ux) + uc) ** 2, 2: interior(uy) ** 2, 3: interior(uz) ** 2, 4: interior(p) ** 2,
             5: (interior(ux) + uc) * uy_cell(uy), 6: ux_cell(ux, uc) * uy_cell(uy)}
    if scalar:
        prods[9] = interior(sc) ** 2
    for m, pr in prods.items():
        want = pr.mean(axis=(0, 2))
        assert np.allclose(sx[:, :, m - 1].sum(axis=0), want, rtol=1e-13, atol=1e-13), m
        assert np.allclose(sz[:, :, m - 1].sum(axis=0), want, rtol=1e-13, atol=1e-13), m
    assert sx.shape == (9, 12, 9 if scalar else 8) and sz.shape == (6, 12, 9 if scalar else 8)


@pytest.mark.parametrize("half", [False, True])
def test_lcs_single_modes(half):
    nx, ny, nz, m, q = 16, 12, 10, 3, 2
    case = channel(nx, ny, nz, half=half, ucrf=0.25)
    o = O.Oracle(case)
    r = np.random.default_rng(3)
    A = 0.5 + r.random(ny + 2)
    ph = 2 * math.pi * r.random(ny + 2)
    i = np.arange(nx)[:, None, None]
    k = np.arange(nz)[None, None, :]
    core_x = A[None, :, None] * np.cos(2 * math.pi * m * i / nx + ph[None, :, None]) * np.ones((1, 1, nz))
    core_z = A[None, :, None] * np.cos(2 * math.pi * q * k / nz + ph[None, :, None]) * np.ones((nx, 1, 1))
    z = haloed(np.zeros((nx, ny + 2, nz)))
    jl = 3
    for core, d, mode, n in ((core_x, 0, m, nx), (core_z, 1, q, nz)):
        sx, sz = o.spectra_1d(haloed(core), z, z, z, None, jlcs=jl)
        s = (sx, sz)[d]
        for j in range(1, ny + 1):
            jr = jl if (half or j <= ny // 2) else ny + 1 - jl
            dphi = ph[j] - ph[jr]
            e = np.zeros(n // 2 + 1)
            e[mode] = A[j] * A[jr] / 4
            e[0] = case.ucrf ** 2
            assert np.allclose(s[:, j - 1, 6], e * np.cos(np.where(np.arange(n // 2 + 1) == mode, dphi, 0)),
                               atol=1e-14), (d, j)
            want_im = np.zeros(n // 2 + 1)
            want_im[mode] = A[j] * A[jr] * math.sin(dphi) / 4
            assert np.allclose(s[:, j - 1, 7], want_im, atol=1e-14), (d, j)
            # the energy spectrum of the same mode: A^2 / 2 there, uCRF^2 at zero
            en = np.zeros(n // 2 + 1)
            en[mode] = A[j] ** 2 / 2
            en[0] = case.ucrf ** 2
            assert np.allclose(s[:, j - 1, 0], en, atol=1e-14), (d, j)


@pytest.mark.parametrize("half", [False, True])
def test_1d_against_numpy(half):
    nx, ny, nz = 12, 10, 16
    case = channel(nx, ny, nz, scalar=True, ucrf=-0.2, half=half)
    o = O.Oracle(case)
    ux, uy, uz, p, sc = random_fields(case, 11)
    jl = 2
    sx, sz = o.spectra_1d(ux, uy, uz, p, sc, jlcs=jl)
    uc = case.ucrf
    a = {1: interior(ux) + uc, 2: interior(uy), 3: interior(uz), 4: interior(p), 9: interior(sc)}
    pairs = {m: (v, v) for m, v in a.items()}
    pairs[5] = (a[1], uy_cell(uy))
    pairs[6] = (ux_cell(ux, uc), uy_cell(uy))
    for d, (s, n, norm) in enumerate(((sx, nx, 1.0 / (nx * nx * nz)), (sz, nz, 1.0 / (nx * nz * nz)))):
        ax = 0 if d == 0 else 2
        oth = 2 if d == 0 else 0
        w = np.full(n // 2 + 1, 2.0)
        w[0] = w[-1] = 1.0
        for m, (f, g) in pairs.items():
            F, G = np.fft.fft(f, axis=ax), np.fft.fft(g, axis=ax)
            P = (F * G.conj()).real.sum(axis=oth)  # (n, ny) after summing the other direction
            P = P if d == 0 else P.T
            want = P[: n // 2 + 1] * w[:, None] * norm
            assert np.allclose(s[:, :, m - 1], want, rtol=1e-12, atol=1e-13), (d, m)
        F = np.fft.fft(a[1], axis=ax)
        rows = np.array([jl if (half or j <= ny // 2) else ny + 1 - jl for j in range(1, ny + 1)]) - 1
        Fr = F[:, rows, :]
        P = (F * Fr.conj()).sum(axis=oth)
        P = P if d == 0 else P.T
        P = P[: n // 2 + 1]
        im = P.imag.copy()
        im[0] = im[-1] = 0.0
        assert np.allclose(s[:, :, 6], P.real * norm, rtol=1e-12, atol=1e-13), d
        assert np.allclose(s[:, :, 7], im * norm, rtol=1e-12, atol=1e-13), d


@pytest.mark.parametrize("half", [False, True])
def test_2d_against_numpy_half_complex(half):
    nx, ny, nz = 12, 10, 8
    case = channel(nx, ny, nz, scalar=True, ucrf=0.4, half=half)
    o = O.Oracle(case)
    ux, uy, uz, p, sc = random_fields(case, 5)
    rows = (2, 4, 5)
    s2 = o.spectra_2d(ux, uy, uz, p, sc, rows=rows)
    assert s2.shape == (nx, 3, nz, 6)
    norm = (1.0 / (nx * nz)) ** 2 * (1.0 if half else 0.5)
    boxes = {1: interior(ux), 2: interior(uy), 3: interior(uz), 4: interior(p), 6: interior(sc)}
    for m, f in boxes.items():
        H2 = hc(hc(f, 0), 2)
        for q, jct in enumerate(rows):
            v = H2[:, jct - 1, :] ** 2
            if not half:
                jcs = (ny + 2 - jct) if m == 2 else (ny + 1 - jct)
                v = v + H2[:, jcs - 1, :] ** 2
            assert np.allclose(s2[:, q, :, m - 1], norm * v, rtol=1e-12, atol=1e-15), (m, jct)
    Ha = hc(hc(ux_cell(ux, case.ucrf), 0), 2)
    Hb = hc(hc(uy_cell(uy), 0), 2)
    for q, jct in enumerate(rows):
        v = Ha[:, jct - 1, :] * Hb[:, jct - 1, :]
        if not half:
            jcs = ny + 1 - jct
            v = v - Ha[:, jcs - 1, :] * Hb[:, jcs - 1, :]
        assert np.allclose(s2[:, q, :, 4], norm * v, rtol=1e-12, atol=1e-15), jct


def test_samples_accumulate():
    case = channel(8, 8, 8)
    o = O.Oracle(case)
    ux, uy, uz, p = random_fields(case, 2)
    sx, sz = o.spectra_1d(ux, uy, uz, p, jlcs=2)
    sx2, sz2 = o.spectra_1d(ux, uy, uz, p, jlcs=2, specx=sx.copy(order="F"), specz=sz.copy(order="F"))
    assert np.allclose(sx2, 2 * sx, rtol=1e-15) and np.allclose(sz2, 2 * sz, rtol=1e-15)
    s2 = o.spectra_2d(ux, uy, uz, p, rows=(3,))
    s22 = o.spectra_2d(ux, uy, uz, p, rows=(3,), spec2d=s2.copy(order="F"))
    assert np.allclose(s22, 2 * s2, rtol=1e-15)
