"""The reference's statistics files (clcStat, src/m_FlowCase.f90:806-940).

``write_stats`` writes what the reference writes every SaveStat steps from the
accumulated SumStat: a header with the step range and the time-averaged
pressure gradient, then one line per plane j = 1..ny+1 of the NCHASTAT
time-averaged statistics ('(40ES24.15)').  ``stats_name`` is its file name,
Res_Dir/stats<itime as I10.10>.

``write_spectra`` writes the time-averaged energy spectra as the reference
does at the same step: SpecX<itime> and SpecZ<itime> hold EnergySpecX/Z
(nx/2+1 or nz/2+1, ny, NE1) and Spec2D<itime> EnergySpec2D (nx, ny2d, nz, NE2),
each divided by nSpectime and written raw, column-major, in native byte order
(:841-935); ``single`` is the reference's SAVE_SINGLE_Spec2D build option.
"""
from __future__ import annotations

import os

import numpy as np


def stats_name(res_dir: str, itime: int) -> str:
    return os.path.join(res_dir, f"stats{itime:010d}")


def _es(v: float) -> str:
    """Fortran ES24.15: one leading digit, 15 decimals, exponent with sign and two digits."""
    s = f"{v:.15E}"
    mant, exp = s.split("E")
    e = int(exp)
    return f"{mant}E{'+' if e >= 0 else '-'}{abs(e):02d}".rjust(24)


def write_stats(path: str, itime: int, nfstime: int, ivstats: int, is_ux_const: bool, prgradsum: float,
                sumstat: np.ndarray) -> None:
    inf = 1.0 / nfstime
    lines = [f"  The time step range for this fluid statistics is {itime - (nfstime - 1) * ivstats:7d}:"
             f"{ivstats:7d}:{itime:7d}", "  "]
    if is_ux_const:
        lines += ["  Constant velocity in x-dir by adding a pressure gradient.",
                  f"    time averaged pressure gradient is: {_es(abs(prgradsum) * inf)}"]
    else:
        lines += ["  Variable velocity in x-dir while adding a constant body force."]
    lines += ["  "]
    for j in range(sumstat.shape[1]):
        lines.append("".join(_es(v) for v in sumstat[:, j] * inf))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def spectra_name(res_dir: str, kind: str, itime: int) -> str:
    """kind: 'SpecX', 'SpecZ' or 'Spec2D'."""
    return os.path.join(res_dir, f"{kind}{itime:010d}")


def write_spectra(res_dir: str, itime: int, nspectime: int, specx=None, specz=None, spec2d=None,
                  single: bool = False) -> list:
    """Writes the parts that are not None; returns the paths written."""
    inf = 1.0 / nspectime
    out = []
    for kind, arr in (("SpecX", specx), ("SpecZ", specz)):
        if arr is not None:
            path = spectra_name(res_dir, kind, itime)
            np.asarray(arr * inf, dtype=np.float64).ravel(order="F").tofile(path)
            out.append(path)
    if spec2d is not None:
        path = spectra_name(res_dir, "Spec2D", itime)
        np.asarray(spec2d * inf, dtype=np.float32 if single else np.float64).ravel(order="F").tofile(path)
        out.append(path)
    return out
