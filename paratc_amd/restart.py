"""Restart files in the reference's format.

write_restart / read_restart (src/m_IOAndVisu.f90:408-512): raw native-endian
FP64 with no header; each field is the full nx*ny*nz interior in Fortran order
(x fastest): ux + uCRF, uy, uz, pressure, then the scalar for the ScalarFlow
variant, then HistX/Y/Z[/C] only for AB2 with WriteHistOld.  The file name is
RestartDir/RestartFor<RunName><itime as I10.10> (:426).
"""
from __future__ import annotations

import os
from typing import Dict, Optional

import numpy as np


def restart_name(restart_dir: str, run_name: str, itime: int) -> str:
    return os.path.join(restart_dir, f"RestartFor{run_name}{itime:010d}")


def write_restart(path: str, ux, uy, uz, p, scalar=None, hist=None, ucrf: float = 0.0) -> None:
    """Fields are interior (nx, ny, nz) arrays; ux is stored with uCRF added back."""
    with open(path, "wb") as fh:
        for a in [np.asarray(ux) + ucrf, uy, uz, p] + ([scalar] if scalar is not None else []) + list(hist or []):
            np.asfortranarray(a, dtype=np.float64).reshape(-1, order="F").tofile(fh)


def read_restart(path: str, nx: int, ny: int, nz: int, scalar: bool = False, ab2_hist: bool = False,
                 ucrf: float = 0.0) -> Dict[str, Optional[np.ndarray]]:
    """Read and validate the byte count the way read_restart does (:477-493)."""
    n = nx * ny * nz
    nfields = 5 if scalar else 4
    nfull = 9 if scalar else 7
    size = os.path.getsize(path)
    if ab2_hist:
        if size != 8 * n * nfull:
            raise ValueError("file byte wrong1")
    elif size not in (8 * n * nfields, 8 * n * nfull):
        raise ValueError("file byte wrong2")
    raw = np.fromfile(path, dtype=np.float64)
    fields = [raw[q * n:(q + 1) * n].reshape((nx, ny, nz), order="F") for q in range(size // (8 * n))]
    out = {"ux": fields[0] - ucrf, "uy": fields[1], "uz": fields[2], "p": fields[3],
           "scalar": fields[4] if scalar else None}
    if ab2_hist:
        k = 5 if scalar else 4
        out["hist"] = fields[k:]
    return out
