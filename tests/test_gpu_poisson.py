"""GPU parity of the Poisson solve (m_Poisson) against the CPU oracle.

The oracle restates clcPPE_x1 (src/m_Poisson.f90:352-452) with FFTW's R2HC /
HC2R semantics; the HIP path uses complex half spectra.  Both diagonalise the
same operator, so they agree to rounding: the bar is 1e-12 relative (FP64).
"""
import numpy as np
import pytest

from helpers import apply_laplacian, inner, rel_err
from oracle.oracle import Oracle
from paratc_amd import capi, case as CS, ic
from paratc_amd.solver import Solver

pytestmark = pytest.mark.gpu

TOL = 1e-12

# (nx, ny, nz, gamma): radix-2 and the odd radices 3, 5, 7 in x/2 and z
GRIDS = [
    (16, 8, 16, 1.5),
    (32, 24, 16, 2.0),
    (48, 20, 30, 1.8),     # nx/2 = 24 = 8*3, nz = 30 = 2*3*5
    (28, 16, 14, 2.0),     # nx/2 = 14 = 2*7,  nz = 14 = 2*7
    (40, 12, 70, 1.6),     # nx/2 = 20 = 4*5,  nz = 70 = 2*5*7
    (128, 96, 128, 1.8),   # the 128x96x128 channel grid
    (96, 140, 50, 2.2),    # ny = 140 as in the large case
    (32, 192, 16, 1.6),    # ny = 192 as in the gravity current
    (64, 256, 32, 2.0),    # ny = 256 as in the 1024x256x512 solve
]


def solve_both(grid, tri, seed=1, compat=True, rows=0, lines=0, lanes=0, cplx=0):
    nx, ny, nz, gam = grid
    case = CS.case_poisson(nx, ny, nz, gam)
    f = ic.poisson_rhs(case, seed=seed) if compat else np.asfortranarray(
        np.random.default_rng(seed).standard_normal((nx, ny, nz)))
    orc = Oracle(case)
    ref = orc.halo()
    orc.clcPPE(f.copy(order="F"), ref, 1)
    am, ac, ap = (orc.metric(n)[1:ny + 1] for n in ("am2ph", "ac2ph", "ap2ph"))
    rdx2, rdz2 = (nx / case.xlx) ** 2, (nz / case.zlz) ** 2
    lnorm = float((np.abs(am) + np.abs(ac) + np.abs(ap)).max() + (1460 + 2 * 838) / 576 * (rdx2 + rdz2))
    with Solver(case) as s:
        s.poisson_config(capi.PREC_FP64, tri, 0, 0)
        if lines:
            s.poisson_tri_lines(lines)
        if rows:
            s.poisson_chunk_rows(rows)
            cfg = s.poisson_get_config()
            assert cfg["tri"] == "partitioned" and cfg["chunk_rows"] == rows, cfg
            assert cfg["tri_lds_lines"] == lines, cfg
        if lanes:
            s.poisson_tri_lanes(lanes)
            cfg = s.poisson_get_config()
            assert cfg["lanes_per_block"] == lanes and cfg["chunk_rows"] == rows, cfg
        if cplx > 0:
            s.poisson_tri_complex(cplx)
            assert s.poisson_get_knob(capi.POISSON_TRI_COMPLEX) == cplx
            assert s.poisson_get_knob(capi.POISSON_CHUNK_ROWS) == rows
            assert s.poisson_get_knob(capi.POISSON_TRI_LANES) == 0
        elif cplx < 0:
            s.poisson_tri_complex(cplx)
            assert s.poisson_get_knob(capi.POISSON_TRI_COMPLEX) == 0
        s.upload(capi.F_ARR1, f, capi.L_INTERIOR)
        s.clcPPE(capi.F_ARR1, capi.F_HALO)
        out = s.download(capi.F_HALO, capi.L_INTERIOR)
        src_after = s.download(capi.F_ARR1, capi.L_INTERIOR)
        res = s.poisson_residual(capi.F_ARR1, capi.F_HALO)
    return f, inner(ref), out, src_after, res / lnorm


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g[:3])))
@pytest.mark.parametrize("tri", [capi.TRI_PARTITIONED, capi.TRI_THOMAS], ids=["part", "thomas"])
def test_poisson_matches_oracle(grid, tri):
    f, ref, out, src_after, res = solve_both(grid, tri)
    print(f"grid={grid} tri={tri} rel={rel_err(out, ref):.3e} res={res:.3e}", flush=True)
    assert rel_err(out, ref) <= TOL
    # the input is not modified on the device
    assert np.array_equal(src_after, f)
    # pinned (0,0) line: the x-z mean of row j=1 is zero (src/m_Poisson.f90:400-404)
    assert abs(out[:, 0, :].mean()) <= 1e-13 * np.abs(out).max()
    # compatible RHS: residual at round-off relative to |L| |phi| (res is divided by |L|)
    assert res <= 1e-13 * np.abs(out).max()


@pytest.mark.parametrize("ny", [64, 50])
@pytest.mark.parametrize("rows,lanes", [(8, 32), (12, 32), (16, 32), (16, 16)])
def test_partitioned_lane_widths_match_oracle(ny, rows, lanes):
    """Workgroups of 32 and 16 real lines (a wavefront holds several chunks);
    nx = 40 leaves the last workgroup part-filled."""
    f, ref, out, _, res = solve_both((40, ny, 24, 1.9), capi.TRI_PARTITIONED, seed=4, rows=rows, lanes=lanes)
    assert rel_err(out, ref) <= TOL


@pytest.mark.parametrize("ny", [64, 50])
@pytest.mark.parametrize("rows,cplx", [(4, 32), (8, 16), (8, 32), (12, 16), (16, 16), (16, 32)])
def test_partitioned_complex_lines_match_oracle(ny, rows, cplx):
    """One lane per complex line (shared coefficients); nx = 40 gives 21
    complex lines, so the last workgroup is part-filled."""
    f, ref, out, _, res = solve_both((40, ny, 24, 1.9), capi.TRI_PARTITIONED, seed=6, rows=rows, cplx=cplx)
    assert rel_err(out, ref) <= TOL


def test_complex_lines_at_ny_256():
    f, ref, out, _, res = solve_both((64, 256, 32, 2.0), capi.TRI_PARTITIONED, seed=7, rows=16, cplx=16)
    assert rel_err(out, ref) <= TOL


@pytest.mark.parametrize("grid", [(8, 4, 8, 1.5), (96, 20, 12, 1.8), (126, 34, 24, 2.0), (200, 10, 70, 1.6),
                                  (64, 48, 32, 2.0)], ids=lambda g: "x".join(map(str, g[:3])))
def test_residual_matches_the_operator(grid):
    """ptc_poisson_residual = max |f - L phi| with L the operator clcPPE
    inverts.  With f = L phi (numpy) every point's residual is round-off, so a
    wrong neighbour anywhere (x and z wrap, the y walls, partial 64-column
    tiles, nx below one wavefront) shows; with an unrelated f the maximum
    itself is compared."""
    nx, ny, nz, gam = grid
    case = CS.case_poisson(nx, ny, nz, gam)
    orc = Oracle(case)
    rg = np.random.default_rng(nx + ny + nz)
    phi = np.asfortranarray(rg.standard_normal((nx, ny, nz)))
    lphi = np.asfortranarray(apply_laplacian(orc, case, phi))
    f = np.asfortranarray(rg.standard_normal((nx, ny, nz)))
    with Solver(case) as s:
        s.upload(capi.F_HALO, phi, capi.L_INTERIOR)
        s.upload(capi.F_ARR1, lphi, capi.L_INTERIOR)
        r0 = s.poisson_residual(capi.F_ARR1, capi.F_HALO)
        s.upload(capi.F_ARR1, f, capi.L_INTERIOR)
        r1 = s.poisson_residual(capi.F_ARR1, capi.F_HALO)
    scale = float(np.abs(lphi).max())
    ref1 = float(np.abs(f - lphi).max())
    print(grid, f"r0={r0 / scale:.2e} r1={r1:.6e} ref={ref1:.6e}", flush=True)
    assert r0 <= 1e-13 * scale
    assert abs(r1 - ref1) <= 1e-13 * scale


@pytest.mark.parametrize("ny", [64, 60, 50])
@pytest.mark.parametrize("rows", [4, 8, 12, 16])
@pytest.mark.parametrize("lanes", [0, 64], ids=["auto", "real64"])
def test_partitioned_chunk_lengths_match_oracle(ny, rows, lanes):
    """Every compiled chunk length, with and without identity padding rows, in
    the automatic kernel (complex lines) and the 64-lane real-line kernel."""
    f, ref, out, _, res = solve_both((32, ny, 24, 1.9), capi.TRI_PARTITIONED, seed=2, rows=rows, lanes=lanes)
    assert rel_err(out, ref) <= TOL
    assert res <= 1e-13 * np.abs(out).max()


@pytest.mark.parametrize("slab", [1, 5, 8, 24])
def test_slabbed_xz_passes_match_oracle(slab):
    """x/z passes run slab by slab of y planes (ny = 24: slabs of 5 leave a
    ragged last slab; 24 means unblocked)."""
    case = CS.case_poisson(64, 24, 32, 1.8)
    f = ic.poisson_rhs(case, seed=9)
    with Solver(case) as s:
        s.upload(capi.F_ARR1, f, capi.L_INTERIOR)
        s.poisson_config(capi.PREC_FP64, capi.TRI_PARTITIONED, 0, 0)
        s.poisson_slab_planes(slab)
        cfg = s.poisson_get_config()
        assert cfg["slab_planes"] == (slab if slab < 24 else 0), cfg
        s.clcPPE(capi.F_ARR1, capi.F_HALO)
        out = s.download(capi.F_HALO, capi.L_INTERIOR)
        s.poisson_slab_planes(0)
        s.clcPPE(capi.F_ARR1, capi.F_HALO)
        whole = s.download(capi.F_HALO, capi.L_INTERIOR)
    assert np.array_equal(out, whole)
    orc = Oracle(case)
    ref = orc.halo()
    orc.clcPPE(f.copy(order="F"), ref, 1)
    assert rel_err(out, inner(ref)) <= TOL


@pytest.mark.parametrize("lines", [16, 32])
@pytest.mark.parametrize("ny,rows", [(64, 4), (60, 8), (50, 12), (256, 16), (96, 12)])
def test_lds_staged_y_solve_matches_oracle(lines, ny, rows):
    """The LDS-staged partitioned kernel (16 or 32 real lines per workgroup)."""
    grid = (32, ny, 24, 1.9)
    f, ref, out, _, res = solve_both(grid, capi.TRI_PARTITIONED, seed=2, rows=rows, lines=lines)
    assert rel_err(out, ref) <= TOL
    assert res <= 1e-13 * np.abs(out).max()


@pytest.mark.parametrize("grid,ztile", [((64, 8, 512), 16), ((128, 12, 256), 32), ((32, 8, 700), 8),
                                        ((16, 6, 1400), 4)])
def test_wide_z_tiles_match_oracle(grid, ztile):
    nx, ny, nz = grid
    case = CS.case_poisson(nx, ny, nz, 2.0)
    f = ic.poisson_rhs(case, seed=13)
    with Solver(case) as s:
        s.upload(capi.F_ARR1, f, capi.L_INTERIOR)
        s.poisson_config(capi.PREC_FP64, capi.TRI_PARTITIONED, 0, 0)
        s.poisson_fast_ztile(ztile)
        cfg = s.poisson_get_config()
        assert cfg["z_fast"] and cfg["z_tile"] == ztile, cfg
        s.clcPPE(capi.F_ARR1, capi.F_HALO)
        out = s.download(capi.F_HALO, capi.L_INTERIOR)
    orc = Oracle(case)
    ref = orc.halo()
    orc.clcPPE(f.copy(order="F"), ref, 1)
    assert rel_err(out, inner(ref)) <= TOL


def test_automatic_variant_falls_back_to_thomas():
    """ny beyond every partitioned shape (the complex-line kernel reaches 512
    rows: 64 chunks of 8 in 16-line workgroups): AUTO runs Thomas,
    PARTITIONED refuses."""
    case = CS.case_poisson(16, 600, 8, 2.0)
    with Solver(case) as s:
        s.poisson_config(capi.PREC_FP64, capi.TRI_AUTO, 0, 0)
        assert s.poisson_get_config()["tri"] == "thomas"
        with pytest.raises(capi.PtcError, match="partitioned"):
            s.poisson_config(capi.PREC_FP64, capi.TRI_PARTITIONED, 0, 0)
    f, ref, out, _, _ = solve_both((16, 600, 8, 2.0), capi.TRI_AUTO, seed=3)
    assert rel_err(out, ref) <= TOL


def test_tall_lines_take_narrow_workgroups():
    """ny = 272: the automatic plan runs complex lines (34 chunks of 8 rows);
    with complex lines off it is past the 64-lane real shapes and 32-lane
    workgroups of 23 chunks of 12 rows take it."""
    case = CS.case_poisson(16, 272, 8, 2.0)
    with Solver(case) as s:
        s.poisson_config(capi.PREC_FP64, capi.TRI_AUTO, 0, 0)
        assert s.poisson_get_knob(capi.POISSON_TRI_COMPLEX) == 16
        assert s.poisson_get_knob(capi.POISSON_CHUNK_ROWS) == 8
        s.poisson_tri_complex(-1)
        cfg = s.poisson_get_config()
        assert cfg["tri"] == "partitioned" and cfg["lanes_per_block"] == 32 and cfg["chunk_rows"] == 12, cfg
        assert s.poisson_get_knob(capi.POISSON_TRI_COMPLEX) == 0
    f, ref, out, _, _ = solve_both((16, 272, 8, 2.0), capi.TRI_AUTO, seed=3)
    assert rel_err(out, ref) <= TOL
    f, ref, out, _, _ = solve_both((16, 272, 8, 2.0), capi.TRI_AUTO, seed=3, cplx=-1)
    assert rel_err(out, ref) <= TOL


def test_poisson_incompatible_rhs_matches_oracle():
    """Without the mean removed, row j=1 of the (0,0) line absorbs the
    inconsistency in both implementations."""
    f, ref, out, _, _ = solve_both((32, 24, 16, 2.0), capi.TRI_PARTITIONED, seed=5, compat=False)
    assert rel_err(out, ref) <= TOL


@pytest.mark.parametrize("tile", [1, 2, 4, 8, 16])
def test_poisson_z_tiles_agree(tile):
    case = CS.case_poisson(64, 32, 64, 2.0)
    f = ic.poisson_rhs(case, seed=3)
    outs = []
    with Solver(case) as s:
        s.upload(capi.F_ARR1, f, capi.L_INTERIOR)
        for t in (16, tile):
            s.poisson_config(capi.PREC_FP64, capi.TRI_PARTITIONED, t, 0)
            s.clcPPE(capi.F_ARR1, capi.F_HALO)
            outs.append(s.download(capi.F_HALO, capi.L_INTERIOR))
    assert rel_err(outs[1], outs[0]) <= 1e-14


def test_init_poisson_solver_autotunes_and_zeroes():
    case = CS.case_poisson(64, 48, 32, 2.0)
    f = ic.poisson_rhs(case, seed=4)
    with Solver(case) as s:
        s.upload(capi.F_ARR1, f, capi.L_INTERIOR)
        s.InitPoissonSolver(capi.F_ARR1, capi.F_HALO)
        # the search's timings, as the reference prints its choices' times
        log = s.poisson_search_log().splitlines()
        assert sum(ln.startswith("candidate ") for ln in log) >= 10, log
        assert log[-1].startswith("picked ") and log[-1].endswith(" ms"), log[-1]
        # src/m_Poisson.f90:174 -- both arrays are zero afterwards
        assert not s.download(capi.F_ARR1, capi.L_INTERIOR).any()
        assert not s.download(capi.F_HALO, capi.L_INTERIOR).any()
        s.upload(capi.F_ARR1, f, capi.L_INTERIOR)
        s.clcPPE()
        out = s.download(capi.F_HALO, capi.L_INTERIOR)
    orc = Oracle(case)
    ref = orc.halo()
    orc.clcPPE(f.copy(order="F"), ref, 1)
    assert rel_err(out, inner(ref)) <= TOL


def test_poisson_deterministic():
    case = CS.case_poisson(64, 32, 48, 2.0)
    f = ic.poisson_rhs(case, seed=6)
    with Solver(case) as s:
        s.upload(capi.F_ARR1, f, capi.L_INTERIOR)
        s.clcPPE()
        a = s.download(capi.F_HALO, capi.L_INTERIOR)
        s.clcPPE()
        b = s.download(capi.F_HALO, capi.L_INTERIOR)
    assert np.array_equal(a, b)


# lengths with compile-time specialised FFT kernels: x/2 in {8,16,32,48,64,96,128,256,512,768,
# 2304,4608}, z in {16,32,64,96,128,192,256,512,700,1400}
FAST_GRIDS = [(64, 16, 32), (128, 8, 64), (256, 12, 96), (512, 8, 128), (1024, 8, 192),
              (1536, 8, 256), (96, 8, 512), (16, 20, 16), (192, 24, 128),
              (4608, 8, 16), (9216, 6, 16), (16, 8, 700), (32, 6, 1400)]


@pytest.mark.parametrize("grid", FAST_GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_specialised_fft_kernels_match_generic_and_oracle(grid):
    nx, ny, nz = grid
    case = CS.case_poisson(nx, ny, nz, 2.0)
    f = ic.poisson_rhs(case, seed=8)
    with Solver(case) as s:
        s.upload(capi.F_ARR1, f, capi.L_INTERIOR)
        s.poisson_config(capi.PREC_FP64, capi.TRI_PARTITIONED, 0, 0)
        cfg = s.poisson_get_config()
        s.clcPPE(capi.F_ARR1, capi.F_HALO)
        fast = s.download(capi.F_HALO, capi.L_INTERIOR)
        try:  # the generic kernels hold a whole x line in at most 64 KB of LDS
            s.poisson_config(capi.PREC_FP64, capi.TRI_PARTITIONED, 4, 0)
        except capi.PtcError as e:
            assert nx // 2 * 16 > 65536 and "too large" in str(e), e
            generic = None
        else:
            s.clcPPE(capi.F_ARR1, capi.F_HALO)
            generic = s.download(capi.F_HALO, capi.L_INTERIOR)
    assert cfg["x_fast"] and cfg["z_fast"], cfg
    if generic is not None:
        assert rel_err(fast, generic) <= 1e-13
    orc = Oracle(case)
    ref = orc.halo()
    orc.clcPPE(f.copy(order="F"), ref, 1)
    assert rel_err(fast, inner(ref)) <= TOL


@pytest.mark.parametrize("prec", [capi.PREC_FP64, capi.PREC_FP32_REFINE], ids=["fp64", "fp32"])
@pytest.mark.parametrize("grid", [(64, 48, 32), (128, 96, 128), (96, 20, 32)], ids=lambda g: "x".join(map(str, g)))
def test_streaming_passes_match_cached_ones(grid, prec):
    """Nontemporal loads and stores change only the cache policy: the solve is
    bit-identical with streaming forced on and off, and matches the oracle."""
    nx, ny, nz = grid
    case = CS.case_poisson(nx, ny, nz, 2.0)
    f = ic.poisson_rhs(case, seed=10)
    outs = []
    with Solver(case) as s:
        s.upload(capi.F_ARR1, f, capi.L_INTERIOR)
        for mode in (-1, 1):
            s.poisson_config(prec, capi.TRI_AUTO, 0, 1)
            s.poisson_streaming(mode)
            assert s.poisson_get_knob(capi.POISSON_STREAMING) == (1 if mode > 0 else 0)
            s.clcPPE(capi.F_ARR1, capi.F_HALO)
            outs.append(s.download(capi.F_HALO, capi.L_INTERIOR))
        s.poisson_streaming(0)
        assert s.poisson_get_knob(capi.POISSON_STREAMING) == 0  # small spectra stay cached
    assert np.array_equal(outs[0], outs[1])
    orc = Oracle(case)
    ref = orc.halo()
    orc.clcPPE(f.copy(order="F"), ref, 1)
    assert rel_err(outs[1], inner(ref)) <= (TOL if prec == capi.PREC_FP64 else 1e-10)


@pytest.mark.parametrize("prec", [capi.PREC_FP64, capi.PREC_FP32_REFINE], ids=["fp64", "fp32"])
@pytest.mark.parametrize("grid,ztile,slab", [
    ((1024, 6, 32), 0, 0),      # x 512: several lines per workgroup
    ((9216, 6, 16), 0, 0),      # x 4608: one 576-thread line per workgroup
    ((4608, 6, 16), 0, 2),      # x 2304, slabs of two planes
    ((16, 6, 512), 8, 0), ((16, 6, 512), 16, 0),
    ((16, 6, 700), 4, 0), ((32, 6, 700), 8, 0),
    ((16, 6, 1400), 2, 0), ((16, 6, 1400), 4, 0),
    ((1024, 6, 512), 8, 0)], ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else str(v))
def test_persistent_passes_match_one_shot(grid, ztile, slab, prec):
    """The persistent x/z passes that prefetch their next line or column tile
    run the same operations as the one-workgroup-per-line kernels (the
    compiler may fuse different multiply-adds in the two, so they agree to
    rounding rather than bit for bit), and both match the oracle."""
    nx, ny, nz = grid
    case = CS.case_poisson(nx, ny, nz, 2.0)
    f = ic.poisson_rhs(case, seed=12)
    outs = []
    with Solver(case) as s:
        s.upload(capi.F_ARR1, f, capi.L_INTERIOR)
        for mode in (-1, 1, 2, 3):
            # FP32 storage: two correction sweeps (one leaves 2e-10 at the
            # 9216-point rows, either way)
            s.poisson_config(prec, capi.TRI_AUTO, 0, 2)
            if ztile:
                s.poisson_fast_ztile(ztile)
            if slab:
                s.poisson_slab_planes(slab)
            s.poisson_persist(mode)
            got = s.poisson_get_knob(capi.POISSON_PERSIST)
            if mode > 0:
                # x: the 512-, 2304- and 4608-point lengths; z: 1400;
                # mode 2 the z passes only, 3 the x passes only
                want = ((1 if nx // 2 in (512, 2304, 4608) and mode != 2 else 0) |
                        (2 if nz == 1400 and mode != 3 else 0))
                assert got == want, (got, want)
            else:
                assert got == 0
            s.clcPPE(capi.F_ARR1, capi.F_HALO)
            outs.append(s.download(capi.F_HALO, capi.L_INTERIOR))
    # FP32 storage: the two round their FP32 spectra differently before the
    # FP64 correction sweep
    for o in outs[1:]:
        assert rel_err(o, outs[0]) <= (1e-13 if prec == capi.PREC_FP64 else 1e-10)
    orc = Oracle(case)
    ref = orc.halo()
    orc.clcPPE(f.copy(order="F"), ref, 1)
    # FP32 storage: the defect correction stops at the accuracy of the FP64
    # residual, about cond(L) eps; with the 9216-point rows (dx = 4 pi / 9216)
    # that is 2.3e-10 relative, however many sweeps
    for o in outs:
        assert rel_err(o, inner(ref)) <= (TOL if prec == capi.PREC_FP64 else 5e-10)


# ---- FP32 storage + FP64 defect correction ---------------------------------
@pytest.mark.parametrize("grid,slab,lines", [((64, 48, 32, 2.0), 0, 0), ((128, 96, 128, 1.8), 0, 0),
                                             ((128, 96, 128, 1.8), 16, 0), ((96, 140, 64, 2.2), 0, 0),
                                             ((128, 96, 128, 1.8), 0, 16), ((128, 96, 128, 1.8), 0, -16)])
def test_fp32_refinement_converges_to_the_fp64_solution(grid, slab, lines):
    """Spectra stored in FP32 between passes (FP64 arithmetic); each sweep
    adds the solve of the FP64 residual f - L phi."""
    nx, ny, nz, gam = grid
    case = CS.case_poisson(nx, ny, nz, gam)
    f = ic.poisson_rhs(case, seed=12)
    orc = Oracle(case)
    ref = orc.halo()
    orc.clcPPE(f.copy(order="F"), ref, 1)
    ref = inner(ref)
    errs, res = [], []
    with Solver(case) as s:
        s.upload(capi.F_ARR1, f, capi.L_INTERIOR)
        for sweeps in (0, 1, 2, 3):
            s.poisson_config(capi.PREC_FP32_REFINE, capi.TRI_AUTO, 0, sweeps)
            if slab:
                s.poisson_slab_planes(slab)
            if lines > 0:
                s.poisson_tri_lines(lines)
            elif lines < 0:  # the complex-line kernel with -lines lines per workgroup
                s.poisson_tri_complex(-lines)
            cfg = s.poisson_get_config()
            assert cfg["prec"] == "fp32+refine" and cfg["sweeps"] == sweeps, cfg
            s.clcPPE(capi.F_ARR1, capi.F_HALO)
            errs.append(rel_err(s.download(capi.F_HALO, capi.L_INTERIOR), ref))
            res.append(s.poisson_residual(capi.F_ARR1, capi.F_HALO))
        src_after = s.download(capi.F_ARR1, capi.L_INTERIOR)
    print(grid, slab, "errors", [f"{e:.2e}" for e in errs], "residuals", [f"{r:.2e}" for r in res], flush=True)
    assert np.array_equal(src_after, f)
    assert errs[0] < 1e-4
    assert errs[1] < 1e-2 * errs[0] and errs[2] <= max(1e-2 * errs[1], 2e-12)
    assert errs[3] <= 1e-11
    assert res[3] < res[0]


def test_fp32_form_refuses_what_it_cannot_run():
    case = CS.case_poisson(80, 16, 32, 2.0)  # nx/2 = 40 has no specialised kernel
    with Solver(case) as s:
        with pytest.raises(capi.PtcError, match="specialised FFT"):
            s.poisson_config(capi.PREC_FP32_REFINE, capi.TRI_AUTO, 0, 2)
    case = CS.case_poisson(64, 16, 32, 2.0)
    with Solver(case) as s:
        with pytest.raises(capi.PtcError, match="Thomas"):
            s.poisson_config(capi.PREC_FP32_REFINE, capi.TRI_THOMAS, 0, 2)
        s.poisson_config(capi.PREC_FP32_REFINE, capi.TRI_AUTO, 0, 2)
        assert s.poisson_get_config()["tri"] == "partitioned"
