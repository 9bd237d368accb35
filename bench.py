#!/usr/bin/env python3
"""Benchmark of the MI355X hot path (one GPU).

Workload (the metric's headline configuration): the Poisson drop-in alone,
staggered L p = f on 1024 x 256 x 512 (periodic x/z, Neumann y, tanh-stretched
y with gamma = 2.0), f ~ N(0,1) per cell from seed 1 minus its dyp-weighted
mean, FP64.  One step = one clcPPE solve of a resident right-hand side; the
solves alternate between two independent resident right-hand sides (seed 1,
streams 0 and 1).  Inputs are in HBM before the timed region starts.

The JSON line also carries
  * roofline: the dominant pass of the solve (algorithmic bytes 16 B per grid
    point per pass: one FP64 read and one write, complex half spectra) over its
    mean launch time, measured with HIP events on the library's stream;
    traffic is that kernel's HBM bytes per launch from the FETCH_SIZE /
    WRITE_SIZE summary in profiles/ (null if the summary lacks the kernel);
  * cpu_baseline: the CPU oracle (a port of the reference's clcPPE_x1) on the
    same right-hand side, timed on this host's cores over a bounded sample;
  * rk_substep: grid points per second of one RK3 substep of the 128 x 96 x 128
    closed channel (ChannelIterate on the device, fixed dt = 2e-3);
  * ensemble: 64 closed-channel cases of uneven grids served together under
    Poisson arrivals (20/s, 500 steps each): throughput, p50/p99 per-step
    latency, and the time the same cases take one at a time.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpts/s per RK substep + Poisson HBM GB/s; FP32+refine vs FP64; ensemble p50/p99"
HBM_PEAK_GBS = 8000.0
STAGES = ["x_fwd", "z_fwd", "y_solve", "z_inv", "x_inv"]


def stage_kernels(plan: dict, nx: int, nz: int) -> list:
    """Kernel behind each pass for the plan in use, as rocprofv3 names it (the
    name up to the storage type: FP64 spectra).  With slabbed x/z passes the x
    entries carry the x and z kernels together."""
    d = "HIP_vector_type<double"
    h = nx // 2
    st = "true" if plan.get("streaming") else "false"
    persist = plan.get("persist", 0)
    if plan["x_fast"] and persist & 1:
        xf, xi = f"k2p_fft_x_fwd<{h}, {d}, 2u>, {st}>", f"k2p_fft_x_inv<{h}, {d}, 2u>, false, {st}>"
    elif plan["x_fast"]:
        # the one-shot forward pass carries its source form (0: the plain field)
        xf, xi = f"k2_fft_x_fwd<{h}, {d}, 2u>, {st}, 0>", f"k2_fft_x_inv<{h}, {d}, 2u>, false, {st}>"
    else:
        xf, xi = "k_fft_x_fwd<", "k_fft_x_inv<"
    if plan["z_fast"] and persist & 2:
        zf, zi = (f"k2p_fft_z<{nz}, {plan['z_tile']}, {inv}, {d}" for inv in ("false", "true"))
    elif plan["z_fast"]:
        zf, zi = (f"k2_fft_z<{nz}, {plan['z_tile']}, {inv}, {d}" for inv in ("false", "true"))
    else:
        zf, zi = "k_fft_z<" + str(plan["z_tile"]) + ", false>", "k_fft_z<" + str(plan["z_tile"]) + ", true>"
    mm = plan["chunk_rows"]
    if plan["tri"] != "partitioned":
        y = "k_tri_poisson_thomas"
    elif plan.get("tri_complex_lines"):
        y = f"k_tri_poisson_cplx<{mm}, {plan['tri_complex_lines']}, {d}, 2u>, {st}>"
    elif plan.get("tri_lds_lines"):
        y = f"k_tri_poisson_lds<{mm}, {plan['tri_lds_lines']}, {d}"
    else:
        y = f"k_tri_poisson_split<{mm}, {plan['lanes_per_block']}, {d}"
    return [xf, zf, y, zi, xi]


class HipEvents:
    """hipEvent timing on an explicit stream (the library's own stream)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        vp = C.c_void_p
        self.L.hipEventCreate.argtypes = [C.POINTER(vp)]
        self.L.hipEventRecord.argtypes = [vp, vp]
        self.L.hipEventSynchronize.argtypes = [vp]
        self.L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
        self.L.hipEventDestroy.argtypes = [vp]

    def create(self):
        e = C.c_void_p()
        assert self.L.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.L.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, e0, e1):
        assert self.L.hipEventSynchronize(e1) == 0
        t = C.c_float()
        assert self.L.hipEventElapsedTime(C.byref(t), e0, e1) == 0
        return float(t.value)


def progress(msg: str):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def bench_poisson(args, ev):
    from paratc_amd import capi, case as CS
    from paratc_amd.solver import Solver

    case = CS.case_poisson(1024, 256, 512, 2.0)
    n = case.npts
    s = Solver(case)
    # InitPoissonSolver times the variants (tile shape, Thomas vs partitioned)
    # and zeroes both arrays, so it runs before the inputs are made.
    s.InitPoissonSolver(capi.F_ARR1, capi.F_HALO)
    t0 = time.time()
    # the two right-hand sides generated in HBM from the host generator's
    # stream (ic.poisson_rhs(case, seed=1, stream=0 / 1) on the host)
    s.seed_normal(capi.F_ARR1, seed=1, stream=0, compatible=True)
    s.seed_normal(capi.F_WORK1, seed=1, stream=1, compatible=True)
    f1 = s.download(capi.F_ARR1, capi.L_INTERIOR)  # the CPU baseline's input
    t_gen = time.time() - t0
    progress(f"autotuned plan {s.poisson_get_config()}, inputs seeded in HBM ({t_gen:.1f} s with the download)")
    srcs = [capi.F_ARR1, capi.F_WORK1]
    if args.calib_copy:
        # a plain device copy of one field (8 B read + 8 B write per point) to
        # scale the FETCH_SIZE / WRITE_SIZE counters of the Poisson passes
        s.copy(capi.F_WORK2, capi.F_ARR1)
    for w in range(args.warmup):
        s.clcPPE(srcs[w % 2], capi.F_HALO)
    s.synchronize()
    stream = s.stream()
    s.poisson_stage_times(1)  # record per-pass events from here on
    e0, e1 = ev.create(), ev.create()
    ev.record(e0, stream)
    for k in range(args.steps):
        s.clcPPE(srcs[k % 2], capi.F_HALO)
    ev.record(e1, stream)
    total_ms = ev.elapsed_ms(e0, e1)
    stage_ms, nsolves = s.poisson_stage_times(0)
    s.synchronize()
    plan = s.poisson_get_config()
    progress(f"timed {args.steps} solves: {total_ms / args.steps:.4f} ms per solve")
    assert nsolves == args.steps
    # correctness guard on the last solve: FP64 residual relative to |L| |p|
    res = s.poisson_residual(srcs[(args.steps - 1) % 2], capi.F_HALO)
    ms_per = total_ms / args.steps
    per_stage = [m / args.steps for m in stage_ms]
    bytes_stage = 16.0 * n  # one FP64 read + one FP64 write per grid point
    if plan["slab_planes"]:
        # slabbed x/z passes: x_fwd / x_inv carry the x+z time, one kernel
        # per pass stands alone only in the y solve
        dom = 2
    else:
        dom = int(np.argmax(per_stage))
    achieved = bytes_stage / (per_stage[dom] * 1e-3) / 1e9
    solve_gbs = 5 * bytes_stage / (ms_per * 1e-3) / 1e9
    out = {
        "case": case, "solver": s, "f1": f1, "ms_per": ms_per, "per_stage": per_stage, "dom": dom,
        "achieved": achieved, "solve_gbs": solve_gbs, "residual": res, "t_gen": t_gen,
        "plan": plan,
    }
    return out


def bench_fp32_refine(r, ev, steps=20, max_sweeps=4):
    """FP32-storage solve + FP64 defect correction on the same resident
    right-hand side: the fewest sweeps whose field error against the FP64
    solution is <= 1e-9 (relative max), its time per solve and the ratio."""
    from paratc_amd import capi

    s = r["solver"]
    s.clcPPE(capi.F_ARR1, capi.F_HALO)
    p64 = s.download(capi.F_HALO, capi.L_INTERIOR)
    scale = float(np.abs(p64).max())
    plan64 = s.poisson_get_config()
    chosen = None
    for sweeps in range(max_sweeps + 1):
        s.poisson_config(capi.PREC_FP32_REFINE, capi.TRI_AUTO, 0, sweeps)
        if plan64["slab_planes"]:
            s.poisson_slab_planes(plan64["slab_planes"])
        s.clcPPE(capi.F_ARR1, capi.F_HALO)
        err = float(np.abs(s.download(capi.F_HALO, capi.L_INTERIOR) - p64).max()) / scale
        progress(f"fp32+refine, {sweeps} sweeps: field error {err:.2e} vs FP64")
        if err <= 1e-9:
            chosen = (sweeps, err)
            break
    if chosen is None:
        return {"error": "no sweep count reached 1e-9", "last_error": err}
    stream = s.stream()
    s.clcPPE(capi.F_ARR1, capi.F_HALO)
    e0, e1 = ev.create(), ev.create()
    ev.record(e0, stream)
    for _ in range(steps):
        s.clcPPE(capi.F_ARR1, capi.F_HALO)
    ev.record(e1, stream)
    ms = ev.elapsed_ms(e0, e1) / steps
    res = s.poisson_residual(capi.F_ARR1, capi.F_HALO)
    plan32 = s.poisson_get_config()
    # the specialised x pass forms the residual as it loads its lines
    fused = plan32["x_fast"]
    # back to the FP64 plan
    s.poisson_config(capi.PREC_FP64, capi.TRI_PARTITIONED if plan64["tri"] == "partitioned" else capi.TRI_THOMAS,
                     plan64["z_tile"] if not (plan64["x_fast"] and plan64["z_fast"]) else 0, 0)
    return {"sweeps": chosen[0], "field_error_vs_fp64": chosen[1], "ms_per_solve": round(ms, 4),
            "speed_vs_fp64": round(r["ms_per"] / ms, 3), "residual_max": res,
            # first solve 48 B/pt (FP64 in, FP32 spectra, FP64 out); a sweep:
            # the correction with its residual formed in the x pass 64 (f and
            # phi read there, the x-inverse adds into phi), or a separate
            # residual 24 + correction 56
            "residual_in_x_pass": bool(fused),
            "bytes_per_point": 48 + (64 if fused else 80) * chosen[0]}


def bench_spectra(r, ev, samples=5):
    """clcStat's energy spectra at the Poisson workload's grid: 1-D spectra of
    all eight quantities and 2-D spectra on four rows, one sample = what the
    reference computes every ivSpec steps.  Device-resident fields: ux, uz copy
    the first right-hand side, uy and p the second."""
    from paratc_amd import capi

    s, case = r["solver"], r["case"]
    s.copy(capi.F_UX, capi.F_ARR1)
    s.copy(capi.F_UY, capi.F_WORK1)
    s.copy(capi.F_UZ, capi.F_ARR1)
    s.copy(capi.F_PR, capi.F_WORK1)
    rows = (2, 16, 64, 128)
    s.init_spectra(True, 8, True, rows)
    s.clc_spectra()
    s.synchronize()
    stream = s.stream()
    e0, e1 = ev.create(), ev.create()
    ev.record(e0, stream)
    for _ in range(samples):
        s.clc_spectra()
    ev.record(e1, stream)
    ms = ev.elapsed_ms(e0, e1) / samples
    _, _, _, n = s.get_spectra(reset=True)
    assert n == samples + 1
    # per point: the cell-centre uy (8 B read + 8 B write), five FP64 forward
    # transforms (ux, uy, uz, p, cell-centre uy; x and z pass, 8 B read + 8 B
    # write each; the cell-centre ux spectrum is formed from the ux one), and
    # the partial sums reading nine spectra (8 B per point each)
    alg = 16 + 5 * 32 + 9 * 8
    progress(f"spectra 1024x256x512: {ms:.3f} ms per sample")
    return {"workload": "clcStat energy spectra, 1-D (8 quantities) + 2-D (5 quantities, 4 rows), "
                        "1024x256x512 FP64", "ms_per_sample": round(ms, 3),
            "alg_bytes_per_point": alg, "hbm_gbs_alg": round(alg * case.npts / (ms * 1e-3) / 1e9, 1)}


def bench_c4a(ev, solves=10):
    """The large single case's Poisson solve at 9216 x 140 x 1400 (gamma 2.2):
    the right-hand side N(0,1) per cell from seed 4 minus its dyp-weighted
    mean, generated in HBM (ptc_seed_normal, the host generator's stream);
    FP64 with per-pass times for the roofline of its dominant pass, then FP32
    storage with one FP64 correction sweep."""
    from paratc_amd import capi, case as CS
    from paratc_amd.solver import Solver

    case = CS.case_poisson(9216, 140, 1400, 2.2)
    n = case.npts
    s = Solver(case)
    t0 = time.time()
    s.InitPoissonSolver(capi.F_ARR1, capi.F_HALO)
    t_tune = time.time() - t0
    s.seed_normal(capi.F_ARR1, seed=4, stream=0, compatible=True)
    plan64 = s.poisson_get_config()
    progress(f"9216x140x1400: autotuned in {t_tune:.1f} s, plan {plan64}; right-hand side seeded in HBM")
    stream = s.stream()

    def timed(stages=False):
        s.clcPPE(capi.F_ARR1, capi.F_HALO)
        s.synchronize()
        if stages:
            s.poisson_stage_times(1)
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, stream)
        for _ in range(solves):
            s.clcPPE(capi.F_ARR1, capi.F_HALO)
        ev.record(e1, stream)
        ms = ev.elapsed_ms(e0, e1) / solves
        st = None
        if stages:
            st, k = s.poisson_stage_times(0)
            st = [v / k for v in st]
        return ms, st, s.poisson_residual(capi.F_ARR1, capi.F_HALO)

    ms64, st64, r64 = timed(stages=True)
    ms_plain, _, _ = timed()
    progress(f"9216x140x1400 FP64: {ms64:.3f} ms per solve (with pass events), {ms_plain:.3f} ms plain")
    # per-kernel pass times for the roofline: a slabbed plan interleaves the
    # x and z kernels slab by slab, so its pass events time them together;
    # the same kernels are timed over whole fields (slabs off)
    plan_k = plan64
    st_k = st64
    if plan64["slab_planes"]:
        s.poisson_slab_planes(0)
        plan_k = s.poisson_get_config()
        _, st_k, _ = timed(stages=True)
        s.poisson_slab_planes(plan64["slab_planes"])
        progress(f"9216x140x1400 pass times without slabs: {[round(v, 3) for v in st_k]}")
    p64 = s.download(capi.F_HALO, capi.L_INTERIOR)
    scale = float(np.abs(p64).max())
    s.poisson_config(capi.PREC_FP32_REFINE, capi.TRI_AUTO, 0, 1)
    ms32, _, r32 = timed()
    err32 = float(np.abs(s.download(capi.F_HALO, capi.L_INTERIOR) - p64).max()) / scale
    del p64
    progress(f"9216x140x1400 FP32 + 1 sweep: {ms32:.3f} ms per solve, field error {err32:.2e} vs FP64")
    s.close()
    ms = ms_plain
    dom = int(np.argmax(st_k))
    names = stage_kernels(plan_k, case.nx, case.nz)
    return {
        "workload": "Poisson 9216x140x1400, gamma 2.2, N(0,1) seed 4 minus dyp-weighted mean (seeded in HBM)",
        "fp64_ms_per_solve": round(ms, 3), "fp64_mpts_per_s": round(n / (ms * 1e-3) / 1e6, 1),
        "fp64_solves_per_s": round(1e3 / ms, 2),
        "fp64_hbm_gbs_alg": round(80.0 * n / (ms * 1e-3) / 1e9, 1),
        "fp64_frac_of_peak": round(80.0 * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "stage_ms": {k: round(v, 4) for k, v in zip(
            ["xz_fwd", "-", "y_solve", "-", "zx_inv"] if plan64["slab_planes"] else STAGES, st64) if k != "-"},
        "kernel_pass_ms": {k: round(v, 4) for k, v in zip(STAGES, st_k)},
        "roofline": {"bound": "hbm", "kernel": names[dom], "launch_ms": round(st_k[dom], 4),
                     "bytes_per_launch": 16 * n,
                     "achieved": round(16.0 * n / (st_k[dom] * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(16.0 * n / (st_k[dom] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "fp32_refine_1_sweep_ms_per_solve": round(ms32, 3), "fp32_refine_speed_vs_fp64": round(ms / ms32, 3),
        "fp32_refine_field_error_vs_fp64": err32,
        "residual_fp64": r64, "residual_fp32_refine": r32, "plan_fp64": plan64,
    }


def bench_gravity_substep(ev, grid, gamma=1.6, steps=5, fp32=True):
    """RK3 substeps of the gravity current (Re 5000, Sc 1, Ri 1, ws 0.02,
    dt 1e-3) at `grid`, the initial front and noise seeded in HBM
    (ptc_seed_fields, seed 2), the Poisson variant chosen by
    InitPoissonSolver; then the same with FP32 Poisson storage and one FP64
    correction sweep, its field error after the same steps stated against
    the FP64 run."""
    from paratc_amd import capi, case as CS
    from paratc_amd.solver import Solver

    nx, ny, nz = grid
    case = CS.case_gravity_current(nx, ny, nz, gamma=gamma)

    def run(prec):
        s = Solver(case)
        s.seed_fields(capi.SEED_FRONT, seed=2, amp=1e-3)
        s.SetBC_and_UpdateHalo(capi.F_UX, capi.F_UY, capi.F_UZ, capi.F_UY_YM)
        s.SetBC_and_UpdateHalo_pr(capi.F_PR)
        s.SetBC_and_UpdateHalo_scalar(capi.F_SCALAR)
        s.InitPoissonSolver(capi.F_ARR1, capi.F_HALO)
        if prec == "fp32":
            cfg = s.poisson_get_config()
            s.poisson_config(capi.PREC_FP32_REFINE, capi.TRI_AUTO, 0, 1)
            if cfg["slab_planes"]:
                s.poisson_slab_planes(cfg["slab_planes"])
        s.L.ptc_channel_iterate(s.h, 1, None)
        s.synchronize()
        stream = s.stream()
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, stream)
        rc = s.L.ptc_channel_iterate(s.h, steps, None)
        ev.record(e1, stream)
        ms = ev.elapsed_ms(e0, e1) / (3 * steps)
        assert rc == 0, s.L.ptc_last_error()
        mon = s.ChannelIterate(1)
        fields = [s.download(f, capi.L_INTERIOR) for f in (capi.F_UX, capi.F_UY, capi.F_UZ, capi.F_SCALAR)]
        plan = s.poisson_get_config()
        rhs_ms = None
        if prec == "fp64":
            # the dominant kernel of the substep, the row-marching RHS, on its
            # own (this case has no flow-rate forcing, so clcRhsXYZ is that
            # one launch): 5 launches after one untimed
            s.PMcoeUpdate(1, 1e-3)
            rhs = lambda: s.clcRhsXYZ(capi.F_UX, capi.F_UY, capi.F_UZ, capi.F_SCALAR, capi.F_ARR1, capi.F_ARR2,
                                      capi.F_HALO, capi.F_ARRC, capi.F_HISTX, capi.F_HISTY, capi.F_HISTZ,
                                      capi.F_HISTC, capi.F_PR)
            rhs()
            e0, e1 = ev.create(), ev.create()
            ev.record(e0, stream)
            for _ in range(5):
                rhs()
            ev.record(e1, stream)
            rhs_ms = ev.elapsed_ms(e0, e1) / 5
        s.close()
        return ms, mon, fields, plan, rhs_ms

    ms, mon, f64, plan, rhs_ms = run("fp64")
    progress(f"{nx}x{ny}x{nz} gravity current: {ms:.3f} ms per RK substep")
    out = {"workload": f"gravity current {nx}x{ny}x{nz}, RK3, dt 1e-3, FP64, scalar (seeded in HBM)",
           "ms_per_substep": round(ms, 3), "mpts_per_s": round(case.npts / (ms * 1e-3) / 1e6, 1),
           # 432 B per point: RHS with the scalar 136, four CN solves 96, Poisson
           # with the divergence formed in its x pass 96, the update with the
           # scalar ghosts 72 + 32
           "hbm_gbs_alg": round(432.0 * case.npts / (ms * 1e-3) / 1e9, 1),
           "divmax_after": float(mon[1]), "plan": plan}
    # k_rhs_rows moves 136 B per point: ux uy uz p c and the four old
    # right-hand sides read, the four right-hand sides and four new ones written
    rhs_bytes = 136 * case.npts
    out["rhs_roofline"] = {"kernel": "k_rhs_rows<true, ...>", "launch_ms": round(rhs_ms, 4),
                           "bytes_per_launch": rhs_bytes,
                           "achieved": round(rhs_bytes / (rhs_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(rhs_bytes / (rhs_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    progress(f"{nx}x{ny}x{nz} RHS rows alone: {rhs_ms:.3f} ms, frac {out['rhs_roofline']['frac']}")
    if fp32:
        ms32, mon32, f32, _, _ = run("fp32")
        errs = {name: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
                for name, a, b in zip(("ux", "uy", "uz", "scalar"), f32, f64)}
        progress(f"{nx}x{ny}x{nz} gravity current, FP32 Poisson + 1 sweep: {ms32:.3f} ms per substep, "
                 f"field errors {errs}")
        out["fp32_refine_1_sweep"] = {"ms_per_substep": round(ms32, 3), "speed_vs_fp64": round(ms / ms32, 3),
                                      "steps": steps + 2, "field_error_vs_fp64": errs,
                                      "divmax_after": float(mon32[1])}
    del f64
    return out


def copy_ceiling(r, ev, reps=20):
    """Measured copy bandwidth on this part: the library's streaming copy
    (k_copy_nt: nontemporal 16-byte loads and stores, four in flight per
    thread) of one 1024 x 258 x 512 field, read once and written once."""
    from paratc_amd import capi

    s = r["solver"]
    s.copy(capi.F_WORK2, capi.F_ARR1)
    s.synchronize()
    stream = s.stream()
    e0, e1 = ev.create(), ev.create()
    ev.record(e0, stream)
    for _ in range(reps):
        s.copy(capi.F_WORK2, capi.F_ARR1)
    ev.record(e1, stream)
    ms = ev.elapsed_ms(e0, e1) / reps
    nbytes = 2 * 8 * 1024 * 258 * 512
    return {"kernel": "k_copy_nt", "bytes_per_launch": nbytes, "ms": round(ms, 4),
            "gbs": round(nbytes / (ms * 1e-3) / 1e9, 1)}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, f, budget_s=10.0, max_solves=20):
    """The oracle's clcPPE_x1 (a plain-C restatement of the reference's
    Choice-1 solve, FFTs by its own mixed-radix code) on the same right-hand
    side: up to budget_s of solves on all host cores this process may use
    (at most 16, the box's CPU share), then one solve on one core."""
    from oracle.oracle import Oracle

    cores = host_cores()
    orc = Oracle(case, threads=cores)
    out = orc.halo()
    times = []
    t_start = time.time()
    while len(times) < max_solves and (time.time() - t_start) < budget_s:
        src = f.copy(order="F")
        t0 = time.time()
        orc.clcPPE(src, out, 1)
        times.append(time.time() - t0)
        progress(f"cpu baseline solve {len(times)}: {times[-1]:.2f} s on {cores} cores")
    t = float(np.mean(times))
    orc1 = Oracle(case, threads=1)
    src = f.copy(order="F")
    t0 = time.time()
    orc1.clcPPE(src, out, 1)
    t1 = time.time() - t0
    progress(f"cpu baseline, one core: {t1:.2f} s per solve")
    orc.L.orc_set_threads(cores)
    return {
        "value": round(case.npts / t / 1e6, 3),
        "unit": "Mpts/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{len(times)} full FP64 Poisson solves (clcPPE_x1 restatement, OpenMP) at "
                  f"{case.nx}x{case.ny}x{case.nz}, mean {t:.3f} s per solve",
        "one_core_mpts_per_s": round(case.npts / t1 / 1e6, 3),
        "host": {"nproc": os.cpu_count(), "usable_cores": cores, "model": cpu_model()},
    }


def bench_rk_substep(ev, steps=30, warmup=3):
    from paratc_amd import capi, case as CS, ic
    from paratc_amd.solver import Solver

    case = CS.case_channel(128, 96, 128, 1.8)
    s = Solver(case)
    ux, uy, uz = ic.channel_fields(case, seed=0, amp=0.1)
    s.upload(capi.F_UX, ux)
    s.upload(capi.F_UY, uy)
    s.upload(capi.F_UZ, uz)
    s.SetBC_and_UpdateHalo(capi.F_UX, capi.F_UY, capi.F_UZ, capi.F_UY_YM)
    s.InitPoissonSolver(capi.F_ARR1, capi.F_HALO)
    s.L.ptc_channel_iterate(s.h, warmup, None)
    s.synchronize()
    stream = s.stream()
    e0, e1 = ev.create(), ev.create()
    ev.record(e0, stream)
    rc = s.L.ptc_channel_iterate(s.h, steps, None)
    ev.record(e1, stream)
    ms = ev.elapsed_ms(e0, e1)
    assert rc == 0, s.L.ptc_last_error()
    mon = s.ChannelIterate(1)
    nsub = steps * 3
    t_sub = ms / nsub
    progress(f"rk substep 128x96x128: {t_sub:.4f} ms")
    s.close()
    return {
        "workload": "closed channel 128x96x128, RK3, dt 2e-3, FP64",
        "mpts_per_s": round(case.npts / (t_sub * 1e-3) / 1e6, 2),
        "ms_per_substep": round(t_sub, 4),
        # 344 B per point: RHS 104, three CN solves 72, Poisson with the
        # divergence formed in its x pass 96, velocity/pressure update 72
        "hbm_gbs_alg": round(344.0 * case.npts / (t_sub * 1e-3) / 1e9, 1),
        "divmax_after": float(mon[1]),
    }


def bench_ensemble(copy_gbs, steps=500, alone_steps=20, budget_gib=64):
    """The 64-case closed-channel ensemble (seed 3) served under its arrival
    process, against the same cases run alone one after another.  Both rates
    are also given as algorithmic HBM GB/s (344 / 432 B per point per RK
    substep); the measured copy rate over the one-at-a-time rate bounds the
    speed-up any serving scheme could reach on these kernels' bytes."""
    from paratc_amd import ensemble as E

    wl = E.c3_workload()
    t = time.time()
    solvers = [E.make_case(c, 100 + q) for q, c in enumerate(wl.cases)]
    progress(f"ensemble: built {len(solvers)} cases in {time.time() - t:.1f} s")
    # alone: every case in a one-case ensemble (graph replays, no overlap)
    alone = []
    for q, s in enumerate(solvers):
        with E.Ensemble(budget_gib << 30) as ens:
            cid = ens.admit(s)
            ens.step(1)
            ens.sync()
            t = time.perf_counter()
            ens.step(alone_steps)
            ens.sync()
            alone.append((time.perf_counter() - t) / alone_steps)
            solvers[q] = ens.release(cid, keep=True)
    seq_s = steps * float(np.sum(alone))
    seq_bytes = steps * sum(E.step_bytes(c) for c in wl.cases)
    progress(f"ensemble: one case at a time would take {seq_s:.1f} s for {steps} steps each")
    r = E.serve(wl, steps, budget_gib << 30, solvers, log=progress)
    progress(f"ensemble: served in {r['wall_s']:.1f} s, p50 {r['p50_ms']:.3f} ms, p99 {r['p99_ms']:.3f} ms")
    # HBM bytes one step of every case moves, from the committed FETCH_SIZE /
    # WRITE_SIZE totals of this workload (null when absent)
    from paratc_amd import profsum
    counted = profsum.read_value(os.path.join(ROOT, "profiles", "ensemble_c3_hbm_traffic.md"),
                                 "step_bytes_counted")
    alg_step = sum(E.step_bytes(c) for c in wl.cases)
    cnt = None
    if counted:
        gbs = counted * steps / r["wall_s"] / 1e9
        cnt = {"step_bytes_counted": counted, "step_bytes_alg": alg_step,
               "counted_over_alg": round(counted / alg_step, 3), "served_hbm_gbs_counted": round(gbs, 1),
               "served_frac_of_copy": round(gbs / copy_gbs, 3),
               "one_at_a_time_hbm_gbs_counted": round(counted * steps / seq_s / 1e9, 1)}
    return {
        "workload": "64 closed-channel cases (seed 3), nx 64-256, ny 48-128, nz 64-192, half with a settling "
                    "scalar, Poisson arrivals 20/s, 500 steps each, dt 2e-3, FP64",
        "wall_s": round(r["wall_s"], 3),
        "case_steps_per_s": round(r["case_steps_per_s"], 1),
        "mpts_steps_per_s": round(r["mpts_steps_per_s"], 1),
        "step_latency_p50_ms": round(r["p50_ms"], 4),
        "step_latency_p99_ms": round(r["p99_ms"], 4),
        "one_at_a_time_s": round(seq_s, 3),
        "speedup_vs_one_at_a_time": round(seq_s / r["wall_s"], 3),
        "hbm_gbs_alg": round(r["alg_gbs"], 1),
        "one_at_a_time_hbm_gbs_alg": round(seq_bytes / seq_s / 1e9, 1),
        "speedup_ceiling_at_copy_rate": round(copy_gbs / (seq_bytes / seq_s / 1e9), 3),
        "admission_wait_p99_s": round(r["admit_wait_p99_s"], 4),
        "hbm_counted": cnt,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-rk", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-storage + refinement leg")
    ap.add_argument("--no-spectra", action="store_true", help="skip the statistics-spectra leg")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 9216x140x1400 Poisson and the 4608x140x700 and 1536x192x256 RK legs")
    ap.add_argument("--large-legs", default="c4a,c4b,c2",
                    help="large legs to run: c4a (9216x140x1400 Poisson), c4b (4608x140x700 substep), "
                         "c2 (1536x192x256 substep)")
    ap.add_argument("--ensemble-steps", type=int, default=500, help="steps per case of the served ensemble")
    ap.add_argument("--ensemble-alone-steps", type=int, default=20,
                    help="timed steps per case when each case runs alone")
    ap.add_argument("--calib-copy", action="store_true",
                    help="launch one plain field copy before the solves (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the whole domain is resident on a single MI355X "
              "and this build has no multi-GPU path (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)
    ev = HipEvents()
    r = bench_poisson(args, ev)
    case = r["case"]
    fp32 = None if args.no_fp32 else bench_fp32_refine(r, ev)
    spec = None if args.no_spectra else bench_spectra(r, ev)
    cpu = None if args.no_cpu_baseline else cpu_baseline(case, r["f1"])
    rk = None if args.no_rk else bench_rk_substep(ev)
    copy = copy_ceiling(r, ev)
    progress(f"copy kernel: {copy['gbs']} GB/s")
    ens = None if args.no_ensemble else bench_ensemble(copy["gbs"], steps=args.ensemble_steps,
                                                       alone_steps=args.ensemble_alone_steps)
    large = None
    if not args.no_large:
        r["solver"].close()  # free the 1024x256x512 fields before the large case
        legs = set(args.large_legs.split(","))
        large = {}
        if "c4a" in legs:
            large["poisson_9216x140x1400"] = bench_c4a(ev)
        if "c4b" in legs:
            large["rk_substep_4608x140x700"] = bench_gravity_substep(ev, (4608, 140, 700), steps=3,
                                                                     fp32=not args.no_fp32)
        if "c2" in legs:
            large["rk_substep_1536x192x256"] = bench_gravity_substep(ev, (1536, 192, 256), steps=5,
                                                                     fp32=not args.no_fp32)
    n = case.npts
    kernel = stage_kernels(r["plan"], case.nx, case.nz)[r["dom"]]
    # HBM bytes per launch of that kernel from the committed FETCH_SIZE /
    # WRITE_SIZE summary of this workload (null when it does not cover it)
    from paratc_amd import profsum
    traffic = profsum.read_traffic(os.path.join(ROOT, "profiles", "poisson_1024x256x512_hbm_traffic.md"), kernel)
    line = {
        "metric": METRIC,
        "value": round(n / (r["ms_per"] * 1e-3) / 1e6, 1),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "poisson_fp64_1024x256x512", "grid": [1024, 256, 512], "gamma": 2.0,
                   "rhs": "N(0,1) seed 1 minus dyp-weighted mean, 2 resident RHS alternating",
                   "precision": "fp64"},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(r["achieved"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kernel, "bytes_per_launch": 16 * n,
                     "launch_ms": round(r["per_stage"][r["dom"]], 4),
                     "copy_ceiling": copy, "frac_of_copy": round(r["achieved"] / copy["gbs"], 4)},
        "cpu_baseline": cpu,
        "poisson": {"solves_per_s": round(1e3 / r["ms_per"], 2),
                    "hbm_gbs_alg": round(r["solve_gbs"], 1),
                    "frac_of_peak": round(r["solve_gbs"] / HBM_PEAK_GBS, 4),
                    "stage_ms": {k: round(v, 4) for k, v in zip(
                        ["xz_fwd", "-", "y_solve", "-", "zx_inv"] if r["plan"]["slab_planes"] else STAGES,
                        r["per_stage"]) if k != "-"},
                    "residual_max": r["residual"],
                    "plan": r["plan"],
                    "fp32_refine": fp32},
        "rk_substep": rk,
        "spectra": spec,
        "ensemble": ens,
        "large": large,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
