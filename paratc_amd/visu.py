"""Visualisation output in the reference's format (m_IOAndVisu).

``dump_visu`` writes one raw binary file per saved field and time step,
``VisuFor<RunName>_<attr>_<ntime, 10 digits>``: the cell-centre values at
every iskip/jskip/kskip-th point from the first, x fastest, in the save
precision (src/m_IOAndVisu.f90:182-385 with decomp_2d_write_every,
src/m_Decomp2d.f90:1232-1389).  The fields are computed on the device
(``ptc_visu_field``).  ``write_xdmf`` writes the ``VisuFor<RunName>.xmf``
index of InitVisu (src/m_IOAndVisu.f90:30-155) that ParaView or VisIt read.
The vortex criteria are ``Q`` (Clc_Q_vor) and ``lamda2`` (Clc_lamda2), as the
reference names their files.
"""
from __future__ import annotations

import os
from typing import Dict, Iterable, Optional

import numpy as np

from . import capi

# attribute name -> device field id, in dump_visu's write order
FIELDS = {"ux": capi.VISU_UX, "uy": capi.VISU_UY, "uz": capi.VISU_UZ, "pr": capi.VISU_PR,
          "scalar": capi.VISU_SCALAR, "wx": capi.VISU_WX, "wy": capi.VISU_WY, "wz": capi.VISU_WZ,
          "wMag": capi.VISU_WMAG, "Q": capi.VISU_Q, "lamda2": capi.VISU_LAMBDA2}
# InitVisu's attribute order in the index (:136-148)
XDMF_ORDER = ("ux", "uy", "uz", "pr", "wx", "wy", "wz", "wMag", "Q", "lamda2", "scalar")


def visu_file(run_name: str, attr: str, ntime: int) -> str:
    """The file name dump_visu and Write_XDMF_One use (:168, :210)."""
    return f"VisuFor{run_name}_{attr}_{ntime:010d}"


def dump_visu(solver, ntime: int, res_dir: str, run_name: str, save: Iterable[str], skip=(1, 1, 1),
              real_bytes: int = 8) -> Dict[str, str]:
    """Write the saved fields of the solver's current state; returns attr -> path."""
    want = set(save)
    unknown = want - set(FIELDS)
    if unknown:
        raise ValueError(f"no device form of visualisation field(s) {sorted(unknown)}")
    if "scalar" in want and not solver.case.scalar:
        raise ValueError("the scalar is saved only in the scalar build")
    out = {}
    for attr, fid in FIELDS.items():
        if attr not in want:
            continue
        data = solver.visu_field(fid, skip, real_bytes)
        path = os.path.join(res_dir, visu_file(run_name, attr, ntime)) if res_dir else visu_file(run_name, attr, ntime)
        np.asarray(data).ravel(order="F").tofile(path)
        out[attr] = path
    return out


def fortran_e(v: float, w: int = 14, d: int = 7) -> str:
    """Fortran Ew.d: 0.ddddddd mantissa, signed two-digit exponent, right-justified."""
    if v == 0.0:
        body = "0." + "0" * d + "E+00"
    else:
        mant, exp = f"{abs(v):.{d - 1}E}".split("E")
        digits = mant.replace(".", "")
        body = ("-" if v < 0 else "") + "0." + digits + f"E{int(exp) + 1:+03d}"
    return body.rjust(w)


def write_xdmf(case, res_dir: str, run_name: str, ifirst: int, ilast: int, save_visu: int,
               save: Iterable[str], set_type: int = 0, real_bytes: int = 8, path: Optional[str] = None) -> str:
    """InitVisu's index file (src/m_IOAndVisu.f90:60-155): the rectilinear mesh
    (nodes for set_type 0 = cell data, centres for 1 = node data) and one grid
    per saved time step with one binary attribute per saved field."""
    want = set(save)
    nxc, nyc, nzc = case.nx, case.ny, case.nz
    dx, dz = case.xlx / nxc, case.zlz / nzc
    yp = np.asarray(case.yp())  # yp(1..nyp)
    if set_type == 0:
        center = '" Center="Cell">'
        nxo, nyo, nzo = nxc + 1, nyc + 1, nzc + 1
        xs = [(i - 1) * dx for i in range(1, nxo + 1)]
        ys = list(yp[:nyo])
        zs = [(k - 1) * dz for k in range(1, nzo + 1)]
    else:
        center = '" Center="Node">'
        nxo, nyo, nzo = nxc, nyc, nzc
        xs = [(i - 1) * dx + dx * 0.5 for i in range(1, nxo + 1)]
        ys = list(0.5 * (yp[:nyc] + yp[1:nyc + 1]))
        zs = [(k - 1) * dz + dz * 0.5 for k in range(1, nzo + 1)]
    L = ['<?xml version="1.0" ?>', '<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>',
         '<Xdmf xmlns:xi="http://www.w3.org/2001/XInclude" Version="2.0">', '<Domain>',
         f'    <Topology name="TOPO" TopologyType="3DRectMesh" Dimensions="{nzo:5d}{nyo:5d}{nxo:5d}"/>',
         '    <Geometry name="GEO" GeometryType="VXVYVZ">']
    for n, vals in ((nxo, xs), (nyo, ys), (nzo, zs)):
        L.append(f'        <DataItem Format="XML" DataType="Float" Precision="{real_bytes:1d}" Endian="Native" '
                 f'Dimensions="{n:5d}">')
        L.append("        " + "".join(fortran_e(float(v)) for v in vals) + " ")
        L.append("        </DataItem>")
    L.append("    </Geometry>")
    steps = list(range(ifirst - 1, ilast + 1, save_visu))
    nflds = (ilast - ifirst + 1) // save_visu + 1
    L += ['    <Grid Name="TimeSeries" GridType="Collection" CollectionType="Temporal">',
          '        <Time TimeType="List">',
          f'        <DataItem Format="XML" NumberType="Int" Dimensions="{nflds:6d}">',
          "        " + "".join(f"{t:10d}" for t in steps) + "        </DataItem>",
          " ", "       </Time>"]
    for t in steps:
        L += [f'        <Grid Name="T{t:010d}" GridType="Uniform">',
              '            <Topology Reference="/Xdmf/Domain/Topology[1]"/>',
              '            <Geometry Reference="/Xdmf/Domain/Geometry[1]"/>']
        for attr in XDMF_ORDER:
            if attr not in want:
                continue
            L += [f'            <Attribute Name="{attr}{center}',
                  f'                <DataItem Format="Binary" DataType="Float" Precision="{real_bytes:1d}" '
                  f'Endian="Native" Dimensions="{nzc:5d}{nyc:5d}{nxc:5d}">',
                  "                    " + visu_file(run_name, attr, t),
                  "                </DataItem>", "            </Attribute>"]
        L.append("        </Grid>")
    L += ["    </Grid>", "</Domain>", "</Xdmf>"]
    if path is None:
        path = os.path.join(res_dir, f"VisuFor{run_name}.xmf")
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")
    return path
