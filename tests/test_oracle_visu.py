"""Pins of the oracle's dump_visu fields (src/m_IOAndVisu.f90:182-385) and of
the visualisation writer's file formats.

Single Fourier modes fix the cell-centre interpolations and the central
differences of the vorticity components through their discrete symbols; a
profile linear in the cell-centre y fixes the wall-normal difference on a
stretched mesh (it averages the two one-sided slopes, exact for linear
profiles); subsampling keeps every skip-th point from the first.
"""
import math
import os

import numpy as np

from oracle import oracle as O
from paratc_amd import capi, case as CS, visu
from test_oracle_rhs import Grid


def fields(case, g):
    z = g.zero()
    return z, z.copy(), z.copy(), z.copy(), z.copy()


def test_centre_interpolations_and_passthrough():
    case = CS.case_channel(16, 12, 10, 1.7)
    case.is_use_crf, case.ucrf_in = True, 0.25
    o = O.Oracle(case)
    g = Grid(case, o)
    kx = 2 * 2 * math.pi / case.xlx
    kz = 3 * 2 * math.pi / case.zlz
    ux = g.field(fx=np.sin(kx * g.xf))
    uz = g.field(fz=np.cos(kz * g.zf))
    uy = g.field(fy=g.yf ** 2)
    p = g.field(fx=np.cos(kx * g.xc), fz=np.sin(kz * g.zc))
    ux_c = o.dump_visu(0, ux, uy, uz, p)
    xc = g.xc[2:-2]
    want = math.cos(kx * g.dx / 2) * np.sin(kx * xc) + case.ucrf
    assert np.allclose(ux_c, want[:, None, None], rtol=0, atol=1e-13)
    uz_c = o.dump_visu(2, ux, uy, uz, p)
    zc = g.zc[2:-2]
    assert np.allclose(uz_c, (math.cos(kz * g.dz / 2) * np.cos(kz * zc))[None, None, :], atol=1e-13)
    uy_c = o.dump_visu(1, ux, uy, uz, p)
    yf = g.yf
    assert np.allclose(uy_c, (0.5 * (yf[1:-1] ** 2 + yf[2:] ** 2))[None, :, None], atol=1e-13)
    assert np.array_equal(o.dump_visu(3, ux, uy, uz, p), p[2:-2, 1:-1, 2:-2])


def test_vorticity_symbols():
    """uy = sin(kx x): wz = dv/dx = sin(kx dx)/dx cos(kx x) (2-cell central
    difference); uz = sin(kx x): wy = -dw/dx likewise; ux = sin(kz z): wy =
    du/dz = sin(kz dz)/dz cos(kz z) (the centre interpolation in x is exact for
    an x-independent field)."""
    case = CS.case_channel(16, 8, 20, 1.5)
    o = O.Oracle(case)
    g = Grid(case, o)
    kx = 3 * 2 * math.pi / case.xlx
    kz = 2 * 2 * math.pi / case.zlz
    z = g.zero()
    uy = g.field(fx=np.sin(kx * g.xc))
    wz = o.dump_visu(6, z, uy, z, z)
    sym = math.sin(kx * g.dx) / g.dx
    assert np.allclose(wz, (sym * np.cos(kx * g.xc[2:-2]))[:, None, None], atol=1e-12)
    assert np.allclose(o.dump_visu(4, z, uy, z, z), 0.0, atol=1e-13)
    uz = g.field(fx=np.sin(kx * g.xc))
    wy = o.dump_visu(5, z, z, uz, z)
    assert np.allclose(wy, -(sym * np.cos(kx * g.xc[2:-2]))[:, None, None], atol=1e-12)
    ux = g.field(fz=np.sin(kz * g.zc))
    wy2 = o.dump_visu(5, ux, z, z, z)
    assert np.allclose(wy2, (math.sin(kz * g.dz) / g.dz * np.cos(kz * g.zc[2:-2]))[None, None, :], atol=1e-12)


def test_wall_normal_difference_exact_on_linear_profiles():
    """ux = a yc (all rows, ghosts included): wz = -du/dy = -a on the stretched
    mesh; uz = b yc: wx = dw/dy = b; |w| = sqrt(a^2 + b^2)."""
    case = CS.case_channel(8, 16, 8, 2.2)
    o = O.Oracle(case)
    g = Grid(case, o)
    a, b = 0.7, -0.3
    z = g.zero()
    ux = g.field(fy=a * g.yc)
    uz = g.field(fy=b * g.yc)
    assert np.allclose(o.dump_visu(6, ux, z, z, z), -a, atol=1e-12)
    assert np.allclose(o.dump_visu(4, z, z, uz, z), b, atol=1e-12)
    assert np.allclose(o.dump_visu(7, ux, z, uz, z), math.hypot(a, b), atol=1e-12)


def test_subsampling_keeps_every_skip_th_point_from_the_first():
    case = CS.case_channel(16, 12, 10, 1.7)
    o = O.Oracle(case)
    rng = np.random.default_rng(4)
    a = [np.asfortranarray(rng.standard_normal((16 + 4, 12 + 2, 10 + 4))) for _ in range(4)]
    for q in range(4):
        o.update_halo(a[q], 2, 2, 2, 2)
    for which in range(8):
        full = o.dump_visu(which, *a)
        sub = o.dump_visu(which, *a, skip=(3, 2, 4))
        assert sub.shape == (6, 6, 3)
        assert np.array_equal(sub, full[::3, ::2, ::4]), which


def test_fortran_e_format():
    assert visu.fortran_e(0.0) == " 0.0000000E+00"
    assert visu.fortran_e(1.0) == " 0.1000000E+01"
    assert visu.fortran_e(-0.00012345678) == "-0.1234568E-03"
    assert visu.fortran_e(3.14159265) == " 0.3141593E+01"


def test_xdmf_index_and_raw_files(tmp_path):
    case = CS.case_channel(8, 6, 4, 1.6)
    path = visu.write_xdmf(case, str(tmp_path), "run", 1, 20, 10, ["ux", "wz", "pr"])
    text = open(path).read().splitlines()
    assert os.path.basename(path) == "VisuForrun.xmf"
    assert text[4] == '    <Topology name="TOPO" TopologyType="3DRectMesh" Dimensions="    5    7    9"/>'
    assert text[7] == "        " + "".join(visu.fortran_e(i * case.xlx / 8) for i in range(9)) + " "
    assert '        <DataItem Format="XML" NumberType="Int" Dimensions="     3">' in text
    assert "        " + "".join(f"{t:10d}" for t in (0, 10, 20)) + "        </DataItem>" in text
    names = [ln.strip() for ln in text if ln.strip().startswith("VisuFor")]
    assert names == [visu.visu_file("run", a, t) for t in (0, 10, 20) for a in ("ux", "pr", "wz")]
    assert names[0] == "VisuForrun_ux_0000000000"
    assert '            <Attribute Name="ux" Center="Cell">' in text

    class FakeSolver:  # the device call's contract: shape and dtype
        def __init__(self, c):
            self.case = c

        def visu_field(self, fid, skip, real_bytes):
            shape = tuple((n + s - 1) // s for n, s in zip((case.nx, case.ny, case.nz), skip))
            return np.full(shape, float(fid), dtype=np.float32 if real_bytes == 4 else np.float64, order="F")

    out = visu.dump_visu(FakeSolver(case), 10, str(tmp_path), "run", ["wz", "ux"], skip=(2, 1, 1), real_bytes=4)
    assert sorted(out) == ["ux", "wz"]
    data = np.fromfile(out["wz"], dtype=np.float32)
    assert data.size == 4 * 6 * 4 and np.all(data == capi.VISU_WZ)
    assert os.path.basename(out["ux"]) == "VisuForrun_ux_0000000010"


def linear_velocity(g, G):
    """ux, uy, uz linear in (x, y, z) with gradient G[i][j] = d u_i / d x_j,
    each at its staggered position, halos included: every difference and
    average of the visualisation stencils is then exact."""
    def lin(a, b, c, fx, fy, fz):
        return g.field(fx=a * fx) + g.field(fy=b * fy) + g.field(fz=c * fz)
    ux = lin(*G[0], g.xf, g.yc, g.zc)
    uy = lin(*G[1], g.xc, g.yf, g.zc)
    uz = lin(*G[2], g.xc, g.yc, g.zf)
    return [np.asfortranarray(a) for a in (ux, uy, uz)]


def test_vortex_criteria_on_linear_fields():
    """Clc_Q_vor and Clc_lamda2 (src/m_Tools.f90:307-442) on uniform velocity
    gradients G on a stretched mesh, against closed forms computed
    independently: Q is the second invariant (1/2)((tr G)^2 - tr(G G)), and
    lambda2 the middle eigenvalue of S^2 + W^2 (S, W the symmetric and
    antisymmetric parts of G) from numpy's symmetric eigensolver, which
    also pins CubicRoot's three-real-roots branch and its ordering."""
    case = CS.case_channel(12, 14, 10, 2.0)
    o = O.Oracle(case)
    g = Grid(case, o)
    rng = np.random.default_rng(7)
    z = g.zero()
    for trial in range(6):
        G = rng.standard_normal((3, 3))
        ux, uy, uz = linear_velocity(g, G)
        q_want = 0.5 * (np.trace(G) ** 2 - np.trace(G @ G))
        S, W = 0.5 * (G + G.T), 0.5 * (G - G.T)
        lam = np.linalg.eigvalsh(S @ S + W @ W)
        q = o.dump_visu(capi.VISU_Q, ux, uy, uz, z)
        l2 = o.dump_visu(capi.VISU_LAMBDA2, ux, uy, uz, z)
        scale = np.abs(G).max() ** 2
        assert np.abs(q - q_want).max() <= 1e-11 * scale, trial
        assert np.abs(l2 - lam[1]).max() <= 1e-10 * scale, (trial, lam)
        # vortex criteria of a skipped subset are the same points
        assert np.array_equal(o.dump_visu(capi.VISU_Q, ux, uy, uz, z, skip=(2, 3, 2)), q[::2, ::3, ::2])


def test_vortex_criteria_of_pure_shear_and_rotation():
    """Pure shear u = a y: S^2 + W^2 = 0, so Q = 0 and lambda2 = 0 through
    CubicRoot's repeated-root branch (t = 0).  Solid rotation u = -w z,
    w = w x about y: S = 0, Q = w^2 > 0 and the eigenvalues of W^2 are
    -w^2, -w^2, 0, so lambda2 = -w^2 (a repeated middle root)."""
    case = CS.case_channel(10, 12, 8, 1.8)
    o = O.Oracle(case)
    g = Grid(case, o)
    z = g.zero()
    a = 0.9
    ux, uy, uz = linear_velocity(g, [[0, a, 0], [0, 0, 0], [0, 0, 0]])
    assert np.abs(o.dump_visu(capi.VISU_Q, ux, uy, uz, z)).max() <= 1e-12
    assert np.abs(o.dump_visu(capi.VISU_LAMBDA2, ux, uy, uz, z)).max() <= 1e-12
    w = 0.6
    ux, uy, uz = linear_velocity(g, [[0, 0, -w], [0, 0, 0], [w, 0, 0]])
    assert np.allclose(o.dump_visu(capi.VISU_Q, ux, uy, uz, z), w * w, rtol=0, atol=1e-12)
    l2 = o.dump_visu(capi.VISU_LAMBDA2, ux, uy, uz, z)
    # a repeated root of a cubic moves by O(sqrt(eps)) under round-off
    assert np.allclose(l2, -w * w, rtol=0, atol=1e-7), (l2.min(), l2.max())
