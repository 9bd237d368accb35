"""The path at the full sizes of the reference workloads, through what does
not depend on size.

- Poisson 1024x256x512 (gamma 2.0, seed 1) and 9216x140x1400 (gamma 2.2,
  seed 4): the FP64 solve satisfies the discrete equation to round-off (device
  residual max |f - L phi| against max |f|), and the FP32-storage form with one
  FP64 defect-correction sweep lands within 1e-9 of it.
- Gravity current 1536x192x256 (gamma 1.6, Re 5000, Ri 1, ws 0.02, dt 1e-3,
  seed 2): one full RK3 step against the oracle at the one-step bar (1e-12;
  p: twice the spread of the reference's own Poisson variants), divergence at
  round-off.
- Gravity current 4608x140x700: ten RK3 steps, the autotuned plan; the
  post-projection divergence stays at round-off and the fields stay finite
  (no oracle: its host arrays alone would need ~60 GB).

Each test prints its timings; they run last so that a slow box shows the
other results first.
"""
import time

import numpy as np
import pytest

from helpers import rel_err
from oracle.oracle import Oracle, State
from paratc_amd import capi as F, case as CS, ic
from paratc_amd.solver import Solver
from test_gpu_tscheme import inner_y, upload_state

pytestmark = pytest.mark.gpu


def log(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("grid,gamma,seed", [((1024, 256, 512), 2.0, 1), ((9216, 140, 1400), 2.2, 4)],
                         ids=["1024x256x512", "9216x140x1400"])
def test_poisson_full_size(grid, gamma, seed):
    case = CS.case_poisson(*grid, gamma)
    t0 = time.time()
    f = ic.poisson_rhs(case, seed=seed)
    fmax = float(np.abs(f).max())
    log(f"{grid}: right-hand side in {time.time() - t0:.1f} s")
    with Solver(case) as s:
        s.InitPoissonSolver(F.F_ARR1, F.F_HALO)
        s.upload(F.F_ARR1, f, F.L_INTERIOR)
        del f
        plan = s.poisson_get_config()
        log(f"plan {plan}")
        s.clcPPE(F.F_ARR1, F.F_HALO)
        r64 = s.poisson_residual(F.F_ARR1, F.F_HALO)
        p64 = s.download(F.F_HALO, F.L_INTERIOR)
        s.poisson_config(F.PREC_FP32_REFINE, F.TRI_AUTO, 0, 1)
        s.clcPPE(F.F_ARR1, F.F_HALO)
        r32 = s.poisson_residual(F.F_ARR1, F.F_HALO)
        p32 = s.download(F.F_HALO, F.L_INTERIOR)
    scale = float(np.abs(p64).max())
    err = float(np.abs(p32 - p64).max()) / scale
    log(f"{grid}: FP64 residual {r64:.2e} (max|f| {fmax:.2f}), FP32+1 sweep residual {r32:.2e}, "
        f"field error vs FP64 {err:.2e}, max|phi| {scale:.3e}")
    assert np.isfinite(p64).all()
    assert r64 <= 1e-9 * fmax
    assert r32 <= 1e-9 * fmax
    assert err <= 1e-9


def gravity_state(case, orc, seed):
    st = State(orc)
    st.ux, st.uy, st.uz, st.scalar = ic.gravity_current_fields(case, seed=seed, amp=1e-3)
    orc.SetBC_and_UpdateHalo(st.ux, st.uy, st.uz, st.uy_ym)
    orc.SetBC_and_UpdateHalo_pr(st.p)
    orc.SetBC_and_UpdateHalo_scalar(st.scalar)
    return st


@pytest.mark.timeout(900)
def test_gravity_current_1536x192x256_one_step():
    case = CS.case_gravity_current(1536, 192, 256, gamma=1.6)
    orc = Oracle(case)
    t0 = time.time()
    st = gravity_state(case, orc, seed=2)
    log(f"initial fields in {time.time() - t0:.1f} s")
    with Solver(case) as s:
        upload_state(s, st.ux, st.uy, st.uz, st.p, st.scalar)
        s.InitPoissonSolver(F.F_ARR1, F.F_HALO)
        mon_g = s.ChannelIterate(1)
        out = {fid: s.download(fid) for fid in (F.F_UX, F.F_UY, F.F_UZ, F.F_PR, F.F_SCALAR)}
    log(f"GPU step done, monitors {mon_g}")
    t0 = time.time()
    mon_o = orc.ChannelIterate(st, 1)
    log(f"oracle step (Choice-1) in {time.time() - t0:.1f} s")
    ref = {F.F_UX: st.ux, F.F_UY: st.uy, F.F_UZ: st.uz, F.F_PR: st.p, F.F_SCALAR: st.scalar}
    p1 = st.p.copy(order="F")
    del st
    orc3 = Oracle(case)
    st3 = gravity_state(case, orc3, seed=2)
    orc3.ChannelIterate(st3, 3)
    spread = rel_err(inner_y(st3.p), inner_y(p1))
    del st3
    log(f"reference Choice-1 vs Choice-3 pressure spread {spread:.2e}")
    for fid, r in ref.items():
        e = rel_err(inner_y(out[fid]), inner_y(r))
        bar = 1e-12 if fid != F.F_PR else max(1e-12, 2.0 * spread)
        log(f"field {fid}: {e:.3e} (bar {bar:.1e})")
        assert e <= bar, fid
    assert rel_err(mon_g[1:5], mon_o[1:5]) <= 1e-9
    assert mon_g[1] <= 1e-10


@pytest.mark.timeout(900)
def test_gravity_current_4608x140x700_ten_steps():
    case = CS.case_gravity_current(4608, 140, 700, gamma=1.6)
    t0 = time.time()
    ux, uy, uz, sc = ic.gravity_current_fields(case, seed=2, amp=1e-3)
    log(f"initial fields in {time.time() - t0:.1f} s")
    with Solver(case) as s:
        upload_state(s, ux, uy, uz, np.zeros_like(ux), sc)
        del ux, uy, uz, sc
        s.SetBC_and_UpdateHalo(F.F_UX, F.F_UY, F.F_UZ, F.F_UY_YM)
        s.SetBC_and_UpdateHalo_pr(F.F_PR)
        s.SetBC_and_UpdateHalo_scalar(F.F_SCALAR)
        s.InitPoissonSolver(F.F_ARR1, F.F_HALO)
        log(f"plan {s.poisson_get_config()}")
        t0 = time.time()
        mons = [s.ChannelIterate(1) for _ in range(10)]
        s.synchronize()
        log(f"10 steps in {time.time() - t0:.2f} s; divergence {[f'{m[1]:.1e}' for m in mons]}; "
            f"vmax {mons[-1][2:5]}")
    for m in mons:
        assert np.isfinite(m).all()
        assert m[1] <= 1e-10
    assert 0.0 < mons[-1][2] < 1.0
