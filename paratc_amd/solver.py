"""Host-side mirror of the reference's hot-path interface over the HIP C-ABI.

One ``Solver`` is one case: it stands for the module globals of m_Parameters,
m_MeshAndMetries and m_Variables, and exposes the reference procedures under
their own names (m_Poisson: InitPoissonSolver / clcPPE /
Destory_Poisson_FFT_Plan; m_TScheme: clcRhsXYZ, clcU1Hat, ...; m_BC_and_Halo:
SetBC_and_UpdateHalo*, ...).  Arguments are field ids (``capi.F_*``) naming
device-resident arrays, so callers alias buffers as the reference does.  Errors
raise ``capi.PtcError`` where the reference aborts.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import check


class Solver:
    def __init__(self, case, vel_limit: float = 0.0, div_limit: float = 0.0):
        self.L = capi.load()
        self.case = case
        self.nx, self.ny, self.nz = case.nx, case.ny, case.nz
        g, p = capi.params_from_case(case, vel_limit, div_limit)
        self.yp = np.ascontiguousarray(case.yp(), dtype=np.float64)
        h = C.c_void_p()
        check(self.L.ptc_create(C.byref(g), C.byref(p), _dp(self.yp), C.byref(h)), "ptc_create")
        self.h = h

    @classmethod
    def from_handle(cls, case, handle) -> "Solver":
        """Wrap an existing context (one handed back by an ensemble)."""
        self = cls.__new__(cls)
        self.L = capi.load()
        self.case = case
        self.nx, self.ny, self.nz = case.nx, case.ny, case.nz
        self.yp = np.ascontiguousarray(case.yp(), dtype=np.float64)
        self.h = handle if isinstance(handle, C.c_void_p) else C.c_void_p(handle)
        return self

    def detach(self) -> C.c_void_p:
        """Give up ownership of the context (an ensemble takes it over)."""
        h, self.h = self.h, None
        return h

    def close(self):
        if getattr(self, "h", None):
            self.L.ptc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- transfers ---------------------------------------------------------
    def shape(self, layout):
        if layout == capi.L_HALO:
            return (self.nx + 4, self.ny + 2, self.nz + 4)
        if layout == capi.L_INTERIOR:
            return (self.nx, self.ny, self.nz)
        return (self.nx, self.nz)

    def upload(self, fid: int, arr: np.ndarray, layout: int = capi.L_HALO):
        a = np.asfortranarray(arr, dtype=np.float64)
        assert a.shape == self.shape(layout), (a.shape, self.shape(layout))
        check(self.L.ptc_upload(self.h, fid, _dp(a), layout), "ptc_upload")

    def download(self, fid: int, layout: int = capi.L_HALO) -> np.ndarray:
        a = np.zeros(self.shape(layout), order="F")
        check(self.L.ptc_download(self.h, fid, _dp(a), layout), "ptc_download")
        return a

    def fill(self, fid: int, value: float = 0.0):
        check(self.L.ptc_fill(self.h, fid, value), "ptc_fill")

    def copy(self, dst: int, src: int):
        check(self.L.ptc_copy(self.h, dst, src), "ptc_copy")

    def synchronize(self):
        check(self.L.ptc_synchronize(self.h), "ptc_synchronize")

    def metric(self, which: int) -> np.ndarray:
        out = np.zeros(self.ny + 2)
        check(self.L.ptc_get_metric(self.h, which, _dp(out)), "ptc_get_metric")
        return out

    # ---- seeded inputs generated in HBM -----------------------------------------
    def seed_normal(self, fid: int, seed: int, stream: int = 0, compatible: bool = True):
        """Interior of fid = N(0,1) per cell from (seed, stream), as
        ic.poisson_rhs draws it on the host; compatible: minus its dyp-weighted
        mean (a compatible Neumann right-hand side)."""
        check(self.L.ptc_seed_normal(self.h, fid, seed, stream, 1 if compatible else 0), "ptc_seed_normal")

    def seed_fields(self, kind: int, seed: int, amp: float, x0: float = 2.0, delta: float = 0.05):
        """Initial ux, uy, uz (and the scalar front) as ic.channel_fields
        (kind capi.SEED_CHANNEL) or ic.gravity_current_fields
        (capi.SEED_FRONT) draw them on the host; ghost rows zeroed."""
        check(self.L.ptc_seed_fields(self.h, kind, seed, amp, x0, delta), "ptc_seed_fields")

    # ---- m_Parameters --------------------------------------------------------
    def PMcoeUpdate(self, ns: int, dt: float):
        check(self.L.ptc_pm_coe_update(self.h, ns, dt), "PMcoeUpdate")

    def coe(self) -> np.ndarray:
        o = np.zeros(9)
        check(self.L.ptc_get_coe(self.h, _dp(o)), "ptc_get_coe")
        return o

    def set_coe(self, o):
        o = np.ascontiguousarray(o, dtype=np.float64)
        check(self.L.ptc_set_coe(self.h, _dp(o)), "ptc_set_coe")

    # ---- m_Poisson ----------------------------------------------------------
    def InitPoissonSolver(self, prsrc: int = capi.F_ARR1, prphiHalo: int = capi.F_HALO):
        check(self.L.ptc_init_poisson_solver(self.h, prsrc, prphiHalo), "InitPoissonSolver")

    def clcPPE(self, prsrc: int = capi.F_ARR1, prphiHalo: int = capi.F_HALO):
        check(self.L.ptc_clc_ppe(self.h, prsrc, prphiHalo), "clcPPE")

    def Destory_Poisson_FFT_Plan(self):
        check(self.L.ptc_destroy_poisson_plan(self.h), "Destory_Poisson_FFT_Plan")

    def poisson_config(self, prec=capi.PREC_FP64, tri=capi.TRI_AUTO, tile=0, sweeps=2):
        check(self.L.ptc_poisson_config(self.h, prec, tri, tile, sweeps), "ptc_poisson_config")

    def poisson_chunk_rows(self, rows: int = 0):
        check(self.L.ptc_poisson_set(self.h, capi.POISSON_CHUNK_ROWS, rows), "ptc_poisson_set")

    def poisson_slab_planes(self, planes: int = 0):
        check(self.L.ptc_poisson_set(self.h, capi.POISSON_SLAB_PLANES, planes), "ptc_poisson_set")

    def poisson_fast_ztile(self, tile: int = 0):
        check(self.L.ptc_poisson_set(self.h, capi.POISSON_FAST_ZTILE, tile), "ptc_poisson_set")

    def poisson_tri_lines(self, lines: int = 0):
        check(self.L.ptc_poisson_set(self.h, capi.POISSON_TRI_LINES, lines), "ptc_poisson_set")

    def poisson_tri_lanes(self, lanes: int = 0):
        check(self.L.ptc_poisson_set(self.h, capi.POISSON_TRI_LANES, lanes), "ptc_poisson_set")

    def poisson_tri_complex(self, lines: int = 0):
        check(self.L.ptc_poisson_set(self.h, capi.POISSON_TRI_COMPLEX, lines), "ptc_poisson_set")

    def poisson_streaming(self, mode: int = 0):
        check(self.L.ptc_poisson_set(self.h, capi.POISSON_STREAMING, mode), "ptc_poisson_set")

    def poisson_persist(self, mode: int = 1):
        check(self.L.ptc_poisson_set(self.h, capi.POISSON_PERSIST, mode), "ptc_poisson_set")

    def rhs_shape(self, segment_threads: int = 0, rows: int = 0, xcd_run: int = 0, min_planes: int = 0):
        """Workgroup shape and z chunking of the clcRhsXYZ kernel (0 = automatic)."""
        for knob, v in ((capi.RHS_SEGMENT_THREADS, segment_threads), (capi.RHS_ROWS, rows),
                        (capi.RHS_XCD_RUN, xcd_run), (capi.RHS_MIN_PLANES, min_planes)):
            check(self.L.ptc_rhs_set(self.h, knob, v), "ptc_rhs_set")

    def rhs_get_shape(self) -> dict:
        out = {}
        for name, knob in (("segment_threads", capi.RHS_SEGMENT_THREADS), ("rows", capi.RHS_ROWS),
                           ("xcd_run", capi.RHS_XCD_RUN), ("min_planes", capi.RHS_MIN_PLANES)):
            v = C.c_int()
            check(self.L.ptc_rhs_get(self.h, knob, C.byref(v)), "ptc_rhs_get")
            out[name] = v.value
        return out

    def poisson_get_knob(self, knob: int) -> int:
        v = C.c_int()
        check(self.L.ptc_poisson_get_knob(self.h, knob, C.byref(v)), "ptc_poisson_get_knob")
        return v.value

    def poisson_get_config(self) -> dict:
        o = (C.c_int * 10)()
        check(self.L.ptc_poisson_get_config(self.h, o), "ptc_poisson_get_config")
        tri = {capi.TRI_THOMAS: "thomas", capi.TRI_PARTITIONED: "partitioned"}.get(o[1], str(o[1]))
        return {"prec": "fp64" if o[0] == capi.PREC_FP64 else "fp32+refine", "tri": tri,
                "z_tile": o[2], "sweeps": o[3], "x_fast": bool(o[4] & 1), "z_fast": bool(o[4] & 2),
                "chunk_rows": o[5], "chunks": o[6], "lanes_per_block": o[7], "slab_planes": o[8],
                "tri_lds_lines": o[9], "tri_complex_lines": self.poisson_get_knob(capi.POISSON_TRI_COMPLEX),
                "streaming": self.poisson_get_knob(capi.POISSON_STREAMING),
                "persist": self.poisson_get_knob(capi.POISSON_PERSIST)}

    def poisson_search_log(self) -> str:
        """The last InitPoissonSolver search: each candidate's knobs and ms per solve."""
        n = C.c_int()
        check(self.L.ptc_poisson_search_log(self.h, None, 0, C.byref(n)), "ptc_poisson_search_log")
        buf = C.create_string_buffer(n.value + 1)
        check(self.L.ptc_poisson_search_log(self.h, buf, n.value + 1, None), "ptc_poisson_search_log")
        return buf.value.decode()

    def poisson_residual(self, src: int, dst: int) -> float:
        r = C.c_double()
        check(self.L.ptc_poisson_residual(self.h, src, dst, C.byref(r)), "ptc_poisson_residual")
        return r.value

    def poisson_stage_times(self, enable: int = -1):
        """(summed ms of the five passes, number of solves) since the last reset."""
        ms = (C.c_float * 5)()
        n = C.c_int()
        check(self.L.ptc_poisson_stage_times(self.h, enable, ms, C.byref(n)), "ptc_poisson_stage_times")
        return list(ms), n.value

    def stream(self) -> int:
        st = C.c_void_p()
        check(self.L.ptc_get_stream(self.h, C.byref(st)), "ptc_get_stream")
        return st.value or 0

    # ---- m_TScheme ----------------------------------------------------------
    def InitTimeScheme(self):
        check(self.L.ptc_init_time_scheme(self.h), "InitTimeScheme")

    def clcRhsXYZ(self, ux, uy, uz, scalar, rhsx, rhsy, rhsz, rhsc, histx, histy, histz, histc, pressure):
        check(self.L.ptc_clc_rhs_xyz(self.h, ux, uy, uz, scalar, rhsx, rhsy, rhsz, rhsc, histx, histy,
                                     histz, histc, pressure), "clcRhsXYZ")

    def clcU1Hat(self, ux, rhsx):
        check(self.L.ptc_clc_u1hat(self.h, ux, rhsx), "clcU1Hat")

    def clcU2Hat(self, uy, rhsy, duy_ym):
        check(self.L.ptc_clc_u2hat(self.h, uy, rhsy, duy_ym), "clcU2Hat")

    def clcU3Hat(self, uz, rhsz):
        check(self.L.ptc_clc_u3hat(self.h, uz, rhsz), "clcU3Hat")

    def clcScalar(self, scalar, rhsc):
        check(self.L.ptc_clc_scalar(self.h, scalar, rhsc), "clcScalar")

    def clcPrSrc(self, ux, uy, uz, prsrc) -> float:
        d = C.c_double()
        check(self.L.ptc_clc_prsrc(self.h, ux, uy, uz, prsrc, C.byref(d)), "clcPrSrc")
        return d.value

    def FluidVelUpdate(self, phi, ux, uy, uz):
        check(self.L.ptc_fluid_vel_update(self.h, phi, ux, uy, uz), "FluidVelUpdate")

    def PressureUpdate(self, pressure, phi):
        check(self.L.ptc_pressure_update(self.h, pressure, phi), "PressureUpdate")

    # ---- m_BC_and_Halo ------------------------------------------------------
    def SetBC_and_UpdateHalo(self, ux, uy, uz, uy_ym):
        check(self.L.ptc_set_bc_and_update_halo(self.h, ux, uy, uz, uy_ym), "SetBC_and_UpdateHalo")

    def SetBC_and_UpdateHaloForPrSrc(self, ux, uy, uz, uy_ym):
        check(self.L.ptc_set_bc_and_update_halo_for_prsrc(self.h, ux, uy, uz, uy_ym),
              "SetBC_and_UpdateHaloForPrSrc")

    def SetBC_and_UpdateHalo_pr(self, p):
        check(self.L.ptc_set_bc_and_update_halo_pr(self.h, p), "SetBC_and_UpdateHalo_pr")

    def SetBC_and_UpdateHalo_scalar(self, s):
        check(self.L.ptc_set_bc_and_update_halo_scalar(self.h, s), "SetBC_and_UpdateHalo_scalar")

    def Update_uy_ym(self, uy_ym, duy_ym):
        check(self.L.ptc_update_uy_ym(self.h, uy_ym, duy_ym), "Update_uy_ym")

    def project_initial(self, ux=capi.F_UX, uy=capi.F_UY, uz=capi.F_UZ, uy_ym=capi.F_UY_YM,
                        prsrc=capi.F_ARR1, phi=capi.F_HALO) -> float:
        """Remove the divergence of an initial velocity field by one projection,
        built from the reference's own procedures: SetBC_and_UpdateHaloForPrSrc,
        clcPrSrc, clcPPE, SetBC_and_UpdateHalo_pr, FluidVelUpdate and
        SetBC_and_UpdateHalo (src/m_ChannelSystem.f90:121-134) with alpha dt = 1,
        so u <- u - G L^-1 D u.  The substep coefficients are restored
        afterwards; prsrc and phi are work arrays.  Needs a Poisson plan.
        Returns max|div u| after the projection (CheckDivergence)."""
        save = self.coe()
        c = save.copy()
        c[2] = 1.0  # pmAlpha
        self.set_coe(c)
        try:
            self.SetBC_and_UpdateHaloForPrSrc(ux, uy, uz, uy_ym)
            self.clcPrSrc(ux, uy, uz, prsrc)
            self.clcPPE(prsrc, phi)
            self.SetBC_and_UpdateHalo_pr(phi)
            self.FluidVelUpdate(phi, ux, uy, uz)
            self.SetBC_and_UpdateHalo(ux, uy, uz, uy_ym)
        finally:
            self.set_coe(save)
        return self.CheckDivergence(ux, uy, uz)

    # ---- m_Tools monitors ---------------------------------------------------
    def CalcMaxCFL(self, ux, uy, uz) -> float:
        d = C.c_double()
        check(self.L.ptc_calc_max_cfl(self.h, ux, uy, uz, C.byref(d)), "CalcMaxCFL")
        return d.value

    def CheckDivergence(self, ux, uy, uz) -> float:
        d = C.c_double()
        check(self.L.ptc_check_divergence(self.h, ux, uy, uz, C.byref(d)), "CheckDivergence")
        return d.value

    def CalcVmax(self, ux, uy, uz) -> np.ndarray:
        v = np.zeros(3)
        check(self.L.ptc_calc_vmax(self.h, ux, uy, uz, _dp(v)), "CalcVmax")
        return v

    def CalcUxAver(self, ux) -> float:
        d = C.c_double()
        check(self.L.ptc_calc_ux_aver(self.h, ux, C.byref(d)), "CalcUxAver")
        return d.value

    def ClcMeanValue(self, ux, scalar) -> np.ndarray:
        v = np.zeros(3)
        check(self.L.ptc_clc_mean_value(self.h, ux, scalar, _dp(v)), "ClcMeanValue")
        return v

    # ---- m_ChannelSystem ----------------------------------------------------
    def step_graphs(self, enable: bool = True):
        """Replay a captured step graph in ChannelIterate (fixed dt) or launch every kernel."""
        check(self.L.ptc_step_graphs(self.h, 1 if enable else 0), "ptc_step_graphs")

    def ChannelIterate(self, nsteps: int = 1) -> np.ndarray:
        mon = np.zeros(7)
        check(self.L.ptc_channel_iterate(self.h, nsteps, _dp(mon)), "ChannelIterate")
        return mon

    # ---- m_FlowCase statistics ------------------------------------------------
    def clcStat(self, ux=capi.F_UX, uy=capi.F_UY, uz=capi.F_UZ, pressure=capi.F_PR, scalar=None):
        """Accumulate one statistics sample (clcStat, src/m_FlowCase.f90:283-428)."""
        if scalar is None:
            scalar = capi.F_SCALAR if self.case.scalar else -1
        check(self.L.ptc_clc_stat(self.h, ux, uy, uz, pressure, scalar), "clcStat")

    def get_stat(self, reset: bool = False):
        """(SumStat as (NCHASTAT, ny+1), nfstime, PrGradsum)."""
        ns = 40 if self.case.scalar else 35
        out = np.zeros((ns, self.ny + 1), order="F")
        n, nf, pg = C.c_int(), C.c_int(), C.c_double()
        check(self.L.ptc_get_stat(self.h, _dp(out), C.byref(n), C.byref(nf), C.byref(pg), 1 if reset else 0),
              "ptc_get_stat")
        return out, nf.value, pg.value

    def init_spectra(self, clc_1d: bool = True, jfor_lcs: int = 1, clc_2d: bool = False, iy_spec2d=()):
        """SpectraOptions / Spectra2DOpt of InitStatVar (src/m_FlowCase.f90:140-251);
        zeroes the spectra accumulators."""
        rows = list(iy_spec2d)
        arr = (C.c_int * max(1, len(rows)))(*rows)
        check(self.L.ptc_init_spectra(self.h, 1 if clc_1d else 0, jfor_lcs, 1 if clc_2d else 0, len(rows), arr),
              "InitStatVar spectra")
        self._spectra = (bool(clc_1d), bool(clc_2d) and len(rows) > 0, len(rows))

    def clc_spectra(self, ux=capi.F_UX, uy=capi.F_UY, uz=capi.F_UZ, pressure=capi.F_PR, scalar=None):
        """One sample of clcStat's energy spectra (src/m_FlowCase.f90:430-805)."""
        if scalar is None:
            scalar = capi.F_SCALAR if self.case.scalar else -1
        check(self.L.ptc_clc_spectra(self.h, ux, uy, uz, pressure, scalar), "clcStat spectra")

    def get_spectra(self, reset: bool = False):
        """(EnergySpecX, EnergySpecZ, EnergySpec2D, nSpectime): the accumulated sums,
        shapes (nx/2+1, ny, NE1), (nz/2+1, ny, NE1), (nx, ny2d, nz, NE2) in Fortran
        order (None where that part is off)."""
        on1, on2, ny2 = getattr(self, "_spectra", (False, False, 0))
        ne1, ne2 = (9, 6) if self.case.scalar else (8, 5)
        sx = np.zeros((self.nx // 2 + 1, self.ny, ne1), order="F") if on1 else None
        sz = np.zeros((self.nz // 2 + 1, self.ny, ne1), order="F") if on1 else None
        s2 = np.zeros((self.nx, ny2, self.nz, ne2), order="F") if on2 else None
        n = C.c_int()
        check(self.L.ptc_get_spectra(self.h, _dp(sx) if on1 else None, _dp(sz) if on1 else None,
                                     _dp(s2) if on2 else None, C.byref(n), 1 if reset else 0), "ptc_get_spectra")
        return sx, sz, s2, n.value

    def visu_field(self, which: int, skip=(1, 1, 1), real_bytes: int = 8) -> np.ndarray:
        """dump_visu's cell-centre field `which` (capi.VISU_*) at the kept points,
        shape (ceil(nx/iskip), ceil(ny/jskip), ceil(nz/kskip)) in Fortran order,
        float64 or float32 (src/m_IOAndVisu.f90:182-385)."""
        shape = tuple((n + s - 1) // s for n, s in zip((self.nx, self.ny, self.nz), skip))
        out = np.zeros(shape, dtype=np.float64 if real_bytes == 8 else np.float32, order="F")
        check(self.L.ptc_visu_field(self.h, which, skip[0], skip[1], skip[2], real_bytes,
                                    out.ctypes.data_as(C.c_void_p), out.size), "ptc_visu_field")
        return out

    def scalar(self, which: int) -> float:
        d = C.c_double()
        check(self.L.ptc_get_scalar(self.h, which, C.byref(d)), "ptc_get_scalar")
        return d.value


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))
