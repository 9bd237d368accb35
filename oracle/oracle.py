"""ctypes front-end of the CPU oracle (oracle/ptc_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module.

Arrays use the reference's Fortran layout (numpy order='F'):
  haloed   (nx+4, ny+2, nz+4)  reference indices x -1..nx+2, y 0..ny+1, z -1..nz+2
  interior (nx, ny, nz)
  plane    (nx, nz)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libptc_oracle.so")


class OrcCfg(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("flow_type", C.c_int), ("scalar", C.c_int), ("is_ux_const", C.c_int),
        ("ischeme", C.c_int), ("icfl", C.c_int), ("sbc_opt", C.c_int * 2), ("pad_", C.c_int),
        ("xlx", C.c_double), ("yly", C.c_double), ("zlz", C.c_double),
        ("ubulk", C.c_double), ("xnu", C.c_double), ("ucrf", C.c_double),
        ("dtmax", C.c_double), ("cflc", C.c_double),
        ("gravity", C.c_double * 3),
        ("uxbc", C.c_double * 6), ("uybc", C.c_double * 6), ("uzbc", C.c_double * 6),
        ("sdiff", C.c_double), ("gev", C.c_double * 3), ("fvv", C.c_double * 3),
        ("sbc_val", C.c_double * 4),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB)
        dp = C.POINTER(C.c_double)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(OrcCfg), dp]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_metric.argtypes = [vp, C.c_int, dp]
        L.orc_wavenumbers.argtypes = [vp, dp, dp]
        L.orc_inverse_tridiagonal.argtypes = [dp, dp, dp, dp, C.c_int, C.c_int]
        L.orc_r2hc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.orc_hc2r.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.orc_poisson.argtypes = [vp, C.c_int, dp, dp]
        L.orc_pm_coe_update.argtypes = [vp, C.c_int, C.c_double]
        L.orc_get_coe.argtypes = [vp, dp]
        L.orc_set_coe.argtypes = [vp, dp]
        L.orc_clc_rhs_xyz.argtypes = [vp] + [dp] * 13
        L.orc_clc_u1hat.argtypes = [vp, dp, dp]
        L.orc_clc_u2hat.argtypes = [vp, dp, dp, dp]
        L.orc_clc_u3hat.argtypes = [vp, dp, dp]
        L.orc_clc_scalar.argtypes = [vp, dp, dp]
        L.orc_clc_prsrc.restype = C.c_double
        L.orc_clc_prsrc.argtypes = [vp, dp, dp, dp, dp]
        L.orc_fluid_vel_update.argtypes = [vp, dp, dp, dp, dp]
        L.orc_pressure_update.argtypes = [vp, dp, dp]
        L.orc_update_halo.argtypes = [vp, dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_set_bc_and_update_halo.argtypes = [vp, dp, dp, dp, dp]
        L.orc_set_bc_for_prsrc.argtypes = [vp, dp, dp, dp, dp]
        L.orc_set_bc_pr.argtypes = [vp, dp]
        L.orc_set_bc_scalar.argtypes = [vp, dp]
        L.orc_update_uy_ym.argtypes = [vp, dp, dp]
        L.orc_calc_max_cfl.restype = C.c_double
        L.orc_calc_max_cfl.argtypes = [vp, dp, dp, dp]
        L.orc_check_divergence.restype = C.c_double
        L.orc_check_divergence.argtypes = [vp, dp, dp, dp]
        L.orc_calc_vmax.argtypes = [vp, dp, dp, dp, dp]
        L.orc_calc_ux_aver.restype = C.c_double
        L.orc_calc_ux_aver.argtypes = [vp, dp]
        L.orc_clc_mean_value.argtypes = [vp, dp, dp, dp]
        L.orc_clc_stat.argtypes = [vp] + [dp] * 6
        L.orc_dump_visu.argtypes = [vp, C.c_int] + [dp] * 5 + [C.c_int] * 3 + [dp]
        L.orc_visu_count.restype = C.c_longlong
        L.orc_visu_count.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.orc_channel_iterate.argtypes = [vp, C.c_int] + [dp] * 15 + [dp]
        L.orc_spectra_count.restype = C.c_int
        L.orc_spectra_count.argtypes = [vp, C.c_int]
        L.orc_spectra_1d.argtypes = [vp] + [dp] * 5 + [C.c_int, dp, dp]
        L.orc_spectra_2d.argtypes = [vp] + [dp] * 5 + [C.c_int, C.POINTER(C.c_int), dp]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1), "need Fortran-ordered float64"
    return a.ctypes.data_as(C.POINTER(C.c_double))


METRICS = ["yc", "dyp", "dyc", "rdyp", "rdyc", "yinterp", "am2ph", "ac2ph", "ap2ph",
           "am2c", "ac2c", "ap2c", "am2cForCN", "ac2cForCN", "ap2cForCN",
           "am2p", "ac2p", "ap2p", "am2cSc", "ac2cSc", "ap2cSc", "am2cScCN", "ac2cScCN", "ap2cScCN"]


def cfg_from_case(case) -> OrcCfg:
    case.validate()
    c = OrcCfg()
    c.nx, c.ny, c.nz = case.nx, case.ny, case.nz
    c.flow_type = case.flow_type
    c.scalar = int(bool(case.scalar))
    c.is_ux_const = int(bool(case.is_ux_const))
    c.ischeme = case.ischeme
    c.icfl = case.icfl
    c.sbc_opt[0], c.sbc_opt[1] = case.scalar_bc_option
    c.xlx, c.yly, c.zlz = case.xlx, case.yly, case.zlz
    c.ubulk, c.xnu, c.ucrf = case.ubulk, case.xnu, case.ucrf
    c.dtmax, c.cflc = case.dtmax, case.cflc
    for q in range(3):
        c.gravity[q] = case.gravity[q]
    xb, yb, zb = case.bc_values()
    for q in range(6):
        c.uxbc[q], c.uybc[q], c.uzbc[q] = xb[q], yb[q], zb[q]
    c.sdiff = case.sdiff
    gev, fvv = case.scalar_vectors()
    for q in range(3):
        c.gev[q], c.fvv[q] = gev[q], fvv[q]
    for q in range(4):
        c.sbc_val[q] = case.scalar_bc_values[q]
    return c


class Oracle:
    """One oracle context per case (the module globals of the reference)."""

    def __init__(self, case, yp=None, threads: int = 0):
        self.case = case
        self.L = lib()
        if threads:
            self.L.orc_set_threads(threads)
        self.cfg = cfg_from_case(case)
        self.yp = np.ascontiguousarray(case.yp() if yp is None else yp, dtype=np.float64)
        self.h = self.L.orc_create(C.byref(self.cfg), _p(self.yp))
        self.nx, self.ny, self.nz = case.nx, case.ny, case.nz

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # array factories ------------------------------------------------------
    def halo(self):
        return np.zeros((self.nx + 4, self.ny + 2, self.nz + 4), order="F")

    def interior(self):
        return np.zeros((self.nx, self.ny, self.nz), order="F")

    def plane(self):
        return np.zeros((self.nx, self.nz), order="F")

    @staticmethod
    def inner(h):
        """Interior view (x 1..nx, y 1..ny, z 1..nz) of a haloed array."""
        return h[2:-2, 1:-1, 2:-2]

    # tables ----------------------------------------------------------------
    def metric(self, name: str) -> np.ndarray:
        out = np.zeros(self.ny + 2)
        self.L.orc_get_metric(self.h, METRICS.index(name), _p(out))
        return out

    def wavenumbers(self):
        wx, wz = np.zeros(self.nx), np.zeros(self.nz)
        self.L.orc_wavenumbers(self.h, _p(wx), _p(wz))
        return wx, wz

    # reference procedures ---------------------------------------------------
    def clcPPE(self, prsrc, prphi, choice: int = 1):
        self.L.orc_poisson(self.h, choice, _p(prsrc), _p(prphi))

    def PMcoeUpdate(self, ns: int, dt: float):
        self.L.orc_pm_coe_update(self.h, ns, dt)

    def coe(self):
        o = np.zeros(9)
        self.L.orc_get_coe(self.h, _p(o))
        return o

    def set_coe(self, o):
        o = np.ascontiguousarray(o, dtype=np.float64)
        self.L.orc_set_coe(self.h, _p(o))

    def clcRhsXYZ(self, ux, uy, uz, scalar, rhsx, rhsy, rhsz, rhsc, histx, histy, histz, histc, p):
        self.L.orc_clc_rhs_xyz(self.h, _p(ux), _p(uy), _p(uz), _p(scalar), _p(rhsx), _p(rhsy),
                               _p(rhsz), _p(rhsc), _p(histx), _p(histy), _p(histz), _p(histc), _p(p))

    def clcU1Hat(self, ux, rhsx):
        self.L.orc_clc_u1hat(self.h, _p(ux), _p(rhsx))

    def clcU2Hat(self, uy, rhsy, duy_ym):
        self.L.orc_clc_u2hat(self.h, _p(uy), _p(rhsy), _p(duy_ym))

    def clcU3Hat(self, uz, rhsz):
        self.L.orc_clc_u3hat(self.h, _p(uz), _p(rhsz))

    def clcScalar(self, s, rhsc):
        self.L.orc_clc_scalar(self.h, _p(s), _p(rhsc))

    def clcPrSrc(self, ux, uy, uz, prsrc) -> float:
        return self.L.orc_clc_prsrc(self.h, _p(ux), _p(uy), _p(uz), _p(prsrc))

    def FluidVelUpdate(self, phi, ux, uy, uz):
        self.L.orc_fluid_vel_update(self.h, _p(phi), _p(ux), _p(uy), _p(uz))

    def PressureUpdate(self, p, phi):
        self.L.orc_pressure_update(self.h, _p(p), _p(phi))

    def update_halo(self, a, xmh, xph, zmh, zph):
        self.L.orc_update_halo(self.h, _p(a), xmh, xph, zmh, zph)

    def SetBC_and_UpdateHalo(self, ux, uy, uz, uy_ym):
        self.L.orc_set_bc_and_update_halo(self.h, _p(ux), _p(uy), _p(uz), _p(uy_ym))

    def SetBC_and_UpdateHaloForPrSrc(self, ux, uy, uz, uy_ym):
        self.L.orc_set_bc_for_prsrc(self.h, _p(ux), _p(uy), _p(uz), _p(uy_ym))

    def SetBC_and_UpdateHalo_pr(self, p):
        self.L.orc_set_bc_pr(self.h, _p(p))

    def SetBC_and_UpdateHalo_scalar(self, s):
        self.L.orc_set_bc_scalar(self.h, _p(s))

    def Update_uy_ym(self, uy_ym, duy_ym):
        self.L.orc_update_uy_ym(self.h, _p(uy_ym), _p(duy_ym))

    def CalcMaxCFL(self, ux, uy, uz) -> float:
        return self.L.orc_calc_max_cfl(self.h, _p(ux), _p(uy), _p(uz))

    def CheckDivergence(self, ux, uy, uz) -> float:
        return self.L.orc_check_divergence(self.h, _p(ux), _p(uy), _p(uz))

    def CalcVmax(self, ux, uy, uz):
        v = np.zeros(3)
        self.L.orc_calc_vmax(self.h, _p(ux), _p(uy), _p(uz), _p(v))
        return v

    def CalcUxAver(self, ux) -> float:
        return self.L.orc_calc_ux_aver(self.h, _p(ux))

    def ClcMeanValue(self, ux, s):
        v = np.zeros(3)
        self.L.orc_clc_mean_value(self.h, _p(ux), _p(s), _p(v))
        return v

    def clcStat(self, ux, uy, uz, p, scalar, sumstat):
        """Adds this sample's plane means to sumstat (NCHASTAT, ny+1), Fortran order."""
        self.L.orc_clc_stat(self.h, _p(ux), _p(uy), _p(uz), _p(p), _p(scalar), _p(sumstat))

    def dump_visu(self, which: int, ux, uy, uz, p, scalar=None, skip=(1, 1, 1)):
        """dump_visu's field `which` (0 ux .. 8 scalar) at the kept cell centres,
        shape (nxo, nyo, nzo) in Fortran order."""
        iskip, jskip, kskip = skip
        shape = tuple((n + s - 1) // s for n, s in zip((self.case.nx, self.case.ny, self.case.nz), skip))
        out = np.zeros(shape, order="F")
        sc = scalar if scalar is not None else ux
        self.L.orc_dump_visu(self.h, which, _p(ux), _p(uy), _p(uz), _p(p), _p(sc), iskip, jskip, kskip, _p(out))
        return out

    def spectra_1d(self, ux, uy, uz, p, scalar=None, jlcs: int = 1, specx=None, specz=None):
        """Adds one sample of clcStat's 1-D spectra to (specx, specz), shapes
        (nx/2+1, ny, NE1) and (nz/2+1, ny, NE1) in Fortran order; returns them."""
        ne1 = 9 if self.case.scalar else 8
        if specx is None:
            specx = np.zeros((self.nx // 2 + 1, self.ny, ne1), order="F")
        if specz is None:
            specz = np.zeros((self.nz // 2 + 1, self.ny, ne1), order="F")
        sc = scalar if scalar is not None else ux
        self.L.orc_spectra_1d(self.h, _p(ux), _p(uy), _p(uz), _p(p), _p(sc), jlcs, _p(specx), _p(specz))
        return specx, specz

    def spectra_2d(self, ux, uy, uz, p, scalar=None, rows=(2,), spec2d=None):
        """Adds one sample of clcStat's 2-D spectra on the 1-based rows to spec2d,
        shape (nx, len(rows), nz, NE2) in Fortran order; returns it."""
        ne2 = 6 if self.case.scalar else 5
        if spec2d is None:
            spec2d = np.zeros((self.nx, len(rows), self.nz, ne2), order="F")
        iy = (C.c_int * len(rows))(*rows)
        sc = scalar if scalar is not None else ux
        self.L.orc_spectra_2d(self.h, _p(ux), _p(uy), _p(uz), _p(p), _p(sc), len(rows), iy, _p(spec2d))
        return spec2d

    def ChannelIterate(self, st: "State", choice: int = 1):
        mon = np.zeros(7)
        self.L.orc_channel_iterate(self.h, choice, _p(st.ux), _p(st.uy), _p(st.uz), _p(st.p),
                                   _p(st.scalar), _p(st.histx), _p(st.histy), _p(st.histz),
                                   _p(st.histc), _p(st.arr1), _p(st.arr2), _p(st.halo),
                                   _p(st.arrc), _p(st.uy_ym), _p(st.duy_ym), _p(mon))
        return mon


class State:
    """The reference's m_Variables arrays for one case (src/m_Variables.f90:10-34)."""

    def __init__(self, orc: Oracle):
        self.ux, self.uy, self.uz, self.p = orc.halo(), orc.halo(), orc.halo(), orc.halo()
        self.histx, self.histy, self.histz = orc.interior(), orc.interior(), orc.interior()
        self.arr1, self.arr2, self.halo = orc.interior(), orc.interior(), orc.halo()
        sc = bool(orc.case.scalar)
        self.scalar = orc.halo() if sc else None
        self.histc = orc.interior() if sc else None
        self.arrc = orc.halo() if sc else None
        self.uy_ym, self.duy_ym = orc.plane(), orc.plane()


def r2hc(x: np.ndarray) -> np.ndarray:
    """FFTW-style R2HC of each column-major line along axis 0 (test helper)."""
    a = np.array(x, dtype=np.float64, order="F", copy=True)
    n = a.shape[0]
    howmany = a.size // n
    lib().orc_r2hc(n, howmany, 1, n, _p(a.reshape(-1, order="F")))
    return a


def hc2r(x: np.ndarray) -> np.ndarray:
    a = np.array(x, dtype=np.float64, order="F", copy=True)
    n = a.shape[0]
    howmany = a.size // n
    lib().orc_hc2r(n, howmany, 1, n, _p(a.reshape(-1, order="F")))
    return a


def inverse_tridiagonal(a, b, c, f):
    """InverseTridiagonal(aj,bj,cj,fj,m,n) on (m, n) Fortran arrays; returns the solution."""
    a, b, c = (np.asfortranarray(v, dtype=np.float64) for v in (a, b, c))
    x = np.array(f, dtype=np.float64, order="F", copy=True)
    m, n = x.shape
    lib().orc_inverse_tridiagonal(_p(a), _p(b), _p(c), _p(x), m, n)
    return x
