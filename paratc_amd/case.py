"""Case description and host-side set-up for the ParaTC hot path.

This module holds what the reference's host reads from its namelists before the
time loop starts (src/m_Parameters.f90:112-188, src/m_MeshAndMetries.f90:143-192)
and passes to the numerics: grid sizes, domain lengths, physical constants,
boundary options and the wall-normal mesh ``yp``.  The compute path lives in the
HIP library; this file only prepares its inputs.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Sequence, Tuple

import numpy as np

FT_CH = 1  # channel (src/m_Parameters.f90:8)
FT_HC = 2  # half channel (src/m_Parameters.f90:9)


@dataclasses.dataclass
class MeshSection:
    """One wall-normal section (namelist MeshOptions, src/m_MeshAndMetries.f90:80)."""

    length: float
    ncells: int
    stret_type: int  # 0 uniform, 1 tanh, 2 sine, 3 geometric
    stret_option: int  # 0 clustered at bottom, 1 at top
    cstret: float


def clc_ycoord(sections: Sequence[MeshSection], yly: float) -> np.ndarray:
    """Node coordinates yp(1..nyp) of the staggered mesh.

    Restates InitMeshAndMetries' section split (src/m_MeshAndMetries.f90:166-181)
    and clcYcoord (src/m_MeshAndMetries.f90:340-408), including the reference's
    use of a default-kind ``real(m)`` in the tanh branch (line 363), which is an
    exact integer conversion for any mesh that fits in memory.
    """
    nsec = len(sections)
    nyc = sum(s.ncells for s in sections)
    yp = np.zeros(nyc + 2)  # index 1..nyp used
    nid = [1] * (nsec + 1)
    ysec = [0.0] * (nsec + 1)
    sumlen = sum(s.length for s in sections)
    for j, s in enumerate(sections):
        nid[j + 1] = nid[j] + s.ncells
        ysec[j + 1] = ysec[j] + s.length / sumlen * yly
    ysec[nsec] = yly
    for j, s in enumerate(sections):
        a, b = nid[j], nid[j + 1]
        y1, y2 = ysec[j], ysec[j + 1]
        m = b - a
        ylen = y2 - y1
        yp[a] = y1
        yp[b] = y2
        if s.stret_type == 0:
            for q in range(m + 1):
                yp[a + q] = y1 + float(q) / float(m) * ylen
        elif s.stret_type == 1:
            tstr = math.tanh(s.cstret)
            for q in range(m + 1):
                xi = float(q) / float(m)
                y0 = math.tanh(s.cstret * (xi - 1.0)) / tstr + 1.0
                if s.stret_option == 0:
                    yp[a + q] = y1 + y0 * ylen
                else:
                    yp[b - q] = y2 - y0 * ylen
        elif s.stret_type == 2:
            tstr = math.sin(0.5 * s.cstret * math.pi)
            for q in range(m + 1):
                xi = float(q) / float(m) - 1.0
                y0 = math.sin(s.cstret * xi * math.pi * 0.5) / tstr + 1.0
                if s.stret_option == 0:
                    yp[a + q] = y1 + y0 * ylen
                else:
                    yp[b - q] = y2 - y0 * ylen
        elif s.stret_type == 3:
            if s.cstret == 1.0:
                for q in range(m + 1):
                    yp[a + q] = y1 + float(q) / float(m) * ylen
            else:
                tstr = (s.cstret ** m - 1.0) / (s.cstret - 1.0)
                for q in range(m):
                    y0 = (s.cstret ** q) / tstr
                    if s.stret_option == 0:
                        yp[a + q + 1] = yp[a + q] + y0 * ylen
                    else:
                        yp[b - q - 1] = yp[b - q] - y0 * ylen
        else:
            raise ValueError("StretType Wrong !!!")
        yp[a] = y1
        yp[b] = y2
    return yp[1 : nyc + 2].copy()


def tanh_channel_sections(ny: int, gamma: float) -> List[MeshSection]:
    """Two mirrored tanh sections: y_j = 1 + tanh(gamma (2 j / ny - 1)) / tanh(gamma)."""
    if ny % 2:
        raise ValueError("channel mesh needs an even ny")
    return [MeshSection(1.0, ny // 2, 1, 0, gamma), MeshSection(1.0, ny // 2, 1, 1, gamma)]


@dataclasses.dataclass
class Case:
    """Everything the hot path reads from the reference's module globals."""

    nx: int
    ny: int
    nz: int
    xlx: float
    yly: float
    zlz: float
    sections: List[MeshSection]
    flow_type: int = FT_CH
    scalar: bool = False  # the -DScalarFlow build
    is_ux_const: bool = True
    ischeme: int = 3
    icfl: int = 2
    ubulk: float = 1.0
    xnu: float = 1.0 / 2800.0
    is_use_crf: bool = False
    ucrf_in: float = 0.0
    dtmax: float = 2e-3
    cflc: float = 1.5
    gravity: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    uxbc: Tuple[float, ...] = (0.0,) * 6
    uybc: Tuple[float, ...] = (0.0,) * 6
    uzbc: Tuple[float, ...] = (0.0,) * 6
    # ScalarFlowOptions (src/m_Parameters.f90:124-129)
    schmidt: float = 1.0
    gravity_eff: float = 0.0
    gravity_dir: Tuple[float, float, float] = (0.0, -1.0, 0.0)
    falling_vel: float = 0.0
    scalar_bc_option: Tuple[int, int] = (-2, -2)
    scalar_bc_values: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 0.0)

    def validate(self) -> None:
        """Input checks of ReadAndInitParameters (src/m_Parameters.f90:147-152)."""
        if self.nx % 2 or self.nz % 2:
            raise ValueError("mod(nxc,2)/=0 || mod(nzc,2)/=0 WRONG")
        if self.flow_type == FT_CH and self.ny % 2:
            raise ValueError("mod(nyc,2)/=0 For channel geomotry. WRONG")
        if sum(s.ncells for s in self.sections) != self.ny:
            raise ValueError("nycSection Wrong 2 !!!")
        if self.ny < 4:
            raise ValueError("ny must be at least 4")
        for n in (self.nx, self.nz):
            m = n
            for f in (2, 3, 5, 7):
                while m % f == 0:
                    m //= f
            if m != 1:
                raise ValueError(f"FFT length {n} must factor into 2, 3, 5 and 7")

    # derived quantities (src/m_Parameters.f90:153-157,176-187)
    @property
    def ucrf(self) -> float:
        return self.ucrf_in if self.is_use_crf else 0.0

    def bc_values(self):
        """Wall values; only uxBcValue is shifted into the moving frame
        (src/m_Parameters.f90:153-157)."""
        s = self.ucrf if self.is_use_crf else 0.0
        return tuple(v - s for v in self.uxbc), tuple(self.uybc), tuple(self.uzbc)

    @property
    def sdiff(self) -> float:
        return self.xnu / self.schmidt

    def scalar_vectors(self):
        gd = np.array(self.gravity_dir, dtype=float)
        nrm = math.sqrt(gd[0] * gd[0] + gd[1] * gd[1] + gd[2] * gd[2])
        gd = gd / nrm if nrm > 1.0e-10 else np.array([0.0, -1.0, 0.0])
        gev = gd * self.gravity_eff
        fvv = gd * self.falling_vel
        fvv[0] = fvv[0] + self.ucrf
        return tuple(float(v) for v in gev), tuple(float(v) for v in fvv)

    def yp(self) -> np.ndarray:
        return clc_ycoord(self.sections, self.yly)

    @property
    def npts(self) -> int:
        return self.nx * self.ny * self.nz


FOURPI = 4.0 * math.pi


def case_channel(nx=128, ny=96, nz=128, gamma=1.8, re_b=2800.0, dt=2e-3) -> Case:
    """Closed turbulent channel, TurbCha0180-style (Lx=4pi, Ly=2, Lz=4pi/3)."""
    return Case(nx=nx, ny=ny, nz=nz, xlx=FOURPI, yly=2.0, zlz=FOURPI / 3.0,
                sections=tanh_channel_sections(ny, gamma), flow_type=FT_CH, scalar=False,
                is_ux_const=True, ischeme=3, icfl=2, ubulk=1.0, xnu=1.0 / re_b, dtmax=dt)


def case_poisson(nx=1024, ny=256, nz=512, gamma=2.0) -> Case:
    """Poisson drop-in alone: periodic x/z, Neumann y, tanh-stretched y."""
    return Case(nx=nx, ny=ny, nz=nz, xlx=FOURPI, yly=2.0, zlz=FOURPI / 3.0,
                sections=tanh_channel_sections(ny, gamma))


def case_gravity_current(nx=1536, ny=192, nz=256, gamma=1.6, re=5000.0, ws=0.02,
                         ri=1.0, dt=1e-3) -> Case:
    """Periodic-x gravity (turbidity) current, domain 24 x 2 x 2, Sc = 1."""
    kappa = 1.0 / re
    return Case(nx=nx, ny=ny, nz=nz, xlx=24.0, yly=2.0, zlz=2.0,
                sections=tanh_channel_sections(ny, gamma), flow_type=FT_CH, scalar=True,
                is_ux_const=False, ischeme=3, icfl=2, ubulk=0.0, xnu=1.0 / re, dtmax=dt,
                schmidt=1.0, gravity_eff=ri, gravity_dir=(0.0, -1.0, 0.0), falling_vel=ws,
                scalar_bc_option=(-2, -3), scalar_bc_values=(0.0, ws / kappa, 0.0, 0.0))
