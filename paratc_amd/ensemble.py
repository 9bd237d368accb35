"""Ensembles of independent cases on one GPU (the serving layer around the path).

``Ensemble`` wraps ptc_ensemble_* of the C-ABI: admission by HBM footprint,
one graph launch per case per step for fixed-dt cases, cases overlapping on
their own streams with no host barrier between rounds, per case-step latency
(from the case's previous step) percentiles.

``c3_workload`` draws the 64-case closed-channel ensemble: with the project
RNG (seed 3) nx from {64, 96, 128, 192, 256}, ny from {48, 64, 96, 128}, nz
from {64, 96, 128, 192}, Re_b uniform in [2000, 5000]; every other case carries
an active settling scalar with ws uniform in [0.005, 0.03]; arrivals form a
Poisson process of rate 20 cases/s.  ``serve`` replays that arrival process
against an ensemble and ``one_at_a_time`` runs the same cases alone, one
after another, for the throughput comparison.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import capi, case as CS, ic, rng
from .capi import PtcError, check
from .solver import Solver

NX_SET = (64, 96, 128, 192, 256)
NY_SET = (48, 64, 96, 128)
NZ_SET = (64, 96, 128, 192)


def case_channel_scalar(nx, ny, nz, gamma=1.8, re_b=2800.0, ws=0.02, dt=2e-3) -> CS.Case:
    """Closed channel at constant flow rate carrying a settling scalar (Sc = 1,
    gravity -y with g_e = 1; no-flux Robin wall on top, settling outflux at the
    bottom, as the gravity-current case)."""
    c = CS.case_channel(nx, ny, nz, gamma, re_b=re_b, dt=dt)
    kappa = c.xnu
    c.scalar = True
    c.schmidt = 1.0
    c.gravity_eff = 1.0
    c.gravity_dir = (0.0, -1.0, 0.0)
    c.falling_vel = ws
    c.scalar_bc_option = (-2, -3)
    c.scalar_bc_values = (0.0, ws / kappa, 0.0, 0.0)
    return c


@dataclasses.dataclass
class Workload:
    cases: List[CS.Case]
    arrivals: np.ndarray  # seconds after the start


def c3_workload(seed: int = 3, ncases: int = 64, rate: float = 20.0) -> Workload:
    u = rng.uniform(seed, 6 * ncases).reshape(ncases, 6)
    cases = []
    for q in range(ncases):
        nx = NX_SET[min(int(u[q, 0] * len(NX_SET)), len(NX_SET) - 1)]
        ny = NY_SET[min(int(u[q, 1] * len(NY_SET)), len(NY_SET) - 1)]
        nz = NZ_SET[min(int(u[q, 2] * len(NZ_SET)), len(NZ_SET) - 1)]
        re_b = 2000.0 + 3000.0 * u[q, 3]
        if q % 2:
            ws = 0.005 + 0.025 * u[q, 4]
            cases.append(case_channel_scalar(nx, ny, nz, re_b=re_b, ws=ws))
        else:
            cases.append(CS.case_channel(nx, ny, nz, 1.8, re_b=re_b))
    gaps = -np.log(np.maximum(rng.uniform(seed, ncases, stream=1), 1e-300)) / rate
    return Workload(cases, np.cumsum(gaps) - gaps[0])


def make_case(case: CS.Case, seed: int) -> Solver:
    """Context with the closed-channel initial field of the 128 x 96 x 128
    workload (u = 1.5 (1 - y^2) plus uniform noise of 0.1 on u, v, w, its
    divergence removed by one projection; a uniform scalar of 1 where the case
    has one), wall ghosts filled and the default Poisson plan (no timing
    search)."""
    s = Solver(case)
    ux, uy, uz = ic.channel_fields(case, seed=seed, amp=0.1)
    s.upload(capi.F_UX, ux)
    s.upload(capi.F_UY, uy)
    s.upload(capi.F_UZ, uz)
    s.SetBC_and_UpdateHalo(capi.F_UX, capi.F_UY, capi.F_UZ, capi.F_UY_YM)
    if case.scalar:
        s.fill(capi.F_SCALAR, 1.0)
        s.SetBC_and_UpdateHalo_scalar(capi.F_SCALAR)
    s.poisson_config(capi.PREC_FP64, capi.TRI_AUTO, 0, 0)
    s.project_initial()
    return s


class Ensemble:
    def __init__(self, budget_bytes: int):
        self.L = capi.load()
        h = C.c_void_p()
        check(self.L.ptc_ensemble_create(C.c_size_t(budget_bytes), C.byref(h)), "ptc_ensemble_create")
        self.h = h
        self.cases: Dict[int, CS.Case] = {}

    def admit(self, solver: Solver) -> Optional[int]:
        """Case id, or None when the case does not fit the HBM budget now."""
        cid = C.c_int()
        rc = self.L.ptc_ensemble_admit(self.h, solver.h, C.byref(cid))
        if rc == capi.PTC_E_LIMIT:
            return None
        check(rc, "ptc_ensemble_admit")
        self.cases[cid.value] = solver.case
        solver.detach()
        return cid.value

    def step(self, nsteps: int = 1):
        """Enqueue nsteps steps of every admitted case (no host barrier)."""
        check(self.L.ptc_ensemble_step(self.h, nsteps), "ptc_ensemble_step")

    def sync(self):
        """Wait for every enqueued step."""
        check(self.L.ptc_ensemble_sync(self.h), "ptc_ensemble_sync")

    def steps_done(self, cid: int) -> int:
        n = C.c_longlong()
        check(self.L.ptc_ensemble_case(self.h, cid, None, C.byref(n)), "ptc_ensemble_case")
        return n.value

    def view(self, cid: int) -> Solver:
        """A non-owning Solver on an admitted case (downloads between steps)."""
        h = C.c_void_p()
        check(self.L.ptc_ensemble_case(self.h, cid, C.byref(h), None), "ptc_ensemble_case")
        return _View.from_handle(self.cases[cid], h)

    def release(self, cid: int, keep: bool = False) -> Optional[Solver]:
        h = C.c_void_p()
        check(self.L.ptc_ensemble_release(self.h, cid, C.byref(h) if keep else None), "ptc_ensemble_release")
        case = self.cases.pop(cid)
        return Solver.from_handle(case, h) if keep else None

    def usage(self):
        used, budget, live = C.c_size_t(), C.c_size_t(), C.c_int()
        check(self.L.ptc_ensemble_usage(self.h, C.byref(used), C.byref(budget), C.byref(live)), "ptc_ensemble_usage")
        return used.value, budget.value, live.value

    def latency(self, reset: bool = False):
        p50, p99, n = C.c_double(), C.c_double(), C.c_longlong()
        check(self.L.ptc_ensemble_latency(self.h, 1 if reset else 0, C.byref(p50), C.byref(p99), C.byref(n)),
              "ptc_ensemble_latency")
        return p50.value, p99.value, n.value

    def close(self):
        if getattr(self, "h", None):
            self.L.ptc_ensemble_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class _View(Solver):
    def close(self):
        self.h = None


# algorithmic HBM bytes per grid point of one RK substep (the FP64 path:
# right-hand side, Crank-Nicolson solves, Poisson with the source formed in
# its x pass, update): 344 without the scalar, 432 with it
SUBSTEP_BYTES = {False: 344.0, True: 432.0}


def step_bytes(case: CS.Case) -> float:
    """Algorithmic bytes of one step (three RK substeps) of a case."""
    return 3.0 * SUBSTEP_BYTES[bool(case.scalar)] * case.npts


def serve(wl: Workload, steps: int, budget_bytes: int, solvers: Sequence[Solver], log=None) -> dict:
    """Replay the arrival process: a case becomes admissible at its arrival
    time, waits in a queue while the HBM budget is full, runs `steps` steps
    alongside every live case, and leaves.  The loop enqueues one step of
    every live case per pass (each case runs ahead on its own stream, bounded
    by the ensemble), so arrivals are admitted while earlier steps run.
    Returns throughput and latency."""
    order = list(np.argsort(wl.arrivals, kind="stable"))
    queue: List[int] = []
    live: Dict[int, int] = {}  # ensemble id -> workload index
    done_pts = 0.0
    done_bytes = 0.0
    case_steps = 0
    admit_wait = []
    with Ensemble(budget_bytes) as ens:
        t0 = time.perf_counter()
        while order or queue or live:
            now = time.perf_counter() - t0
            while order and wl.arrivals[order[0]] <= now:
                queue.append(order.pop(0))
            still = []
            for q in queue:
                cid = ens.admit(solvers[q])
                if cid is None:
                    still.append(q)
                else:
                    live[cid] = q
                    admit_wait.append(time.perf_counter() - t0 - wl.arrivals[q])
            queue = still
            if not live:
                if order:
                    time.sleep(max(0.0, wl.arrivals[order[0]] - (time.perf_counter() - t0)))
                continue
            ens.step(1)
            for cid in list(live):
                q = live[cid]
                case_steps += 1
                done_pts += wl.cases[q].npts
                done_bytes += step_bytes(wl.cases[q])
                if ens.steps_done(cid) >= steps:
                    ens.release(cid)
                    del live[cid]
                    if log:
                        log(f"case {q} ({wl.cases[q].nx}x{wl.cases[q].ny}x{wl.cases[q].nz}) done")
        ens.sync()
        wall = time.perf_counter() - t0
        p50, p99, n = ens.latency()
    return {"wall_s": wall, "case_steps": case_steps, "case_steps_per_s": case_steps / wall,
            "mpts_steps_per_s": done_pts / wall / 1e6, "alg_gbs": done_bytes / wall / 1e9,
            "p50_ms": p50, "p99_ms": p99, "samples": n,
            "admit_wait_p99_s": float(np.percentile(admit_wait, 99)) if admit_wait else 0.0}


def one_at_a_time(solvers: Sequence[Solver], steps: int) -> dict:
    """Each case alone (ptc_channel_iterate, no overlap), timed per step on
    `steps` steps; the 500-step sequential time is extrapolated per case."""
    per_case = []
    for s in solvers:
        s.L.ptc_channel_iterate(s.h, 1, None)  # warm-up
        s.synchronize()
        t = time.perf_counter()
        check(s.L.ptc_channel_iterate(s.h, steps, None), "ptc_channel_iterate")
        s.synchronize()
        per_case.append((time.perf_counter() - t) / steps)
    return {"step_s": per_case}
