"""GPU checks of the ensemble layer (ptc_ensemble_*).

A fixed-dt case is stepped by replaying a captured HIP graph; every replay
must give exactly the fields that ordinary launches give (the kernels and
their arguments are the same and every reduction has a fixed order), also
with other eager work and synchronisations between the replays.
"""
import numpy as np
import pytest

from paratc_amd import capi, case as CS
from paratc_amd import ensemble as E

pytestmark = pytest.mark.gpu

FIELDS = (capi.F_UX, capi.F_UY, capi.F_UZ, capi.F_PR)


def twins(case, seed=4):
    return E.make_case(case, seed), E.make_case(case, seed)


@pytest.mark.parametrize("kind", ["channel", "scalar", "adaptive"])
def test_graph_replays_match_eager_steps(kind):
    if kind == "scalar":
        case = E.case_channel_scalar(32, 24, 16, re_b=3000.0, ws=0.02)
    else:
        case = CS.case_channel(32, 24, 16, 1.8, re_b=3000.0)
        if kind == "adaptive":
            case.icfl = 1
    a, b = twins(case)
    b.step_graphs(False)  # the reference for every replay: ordinary launches
    other = E.make_case(CS.case_channel(16, 16, 16, 1.5), 9)
    fields = FIELDS + ((capi.F_SCALAR,) if case.scalar else ())
    with E.Ensemble(1 << 34) as ens:
        cid = ens.admit(a)
        assert cid is not None
        view = ens.view(cid)
        for step in range(4):
            ens.step(1)  # one replay of the graph captured at admission
            # unrelated eager work and a synchronisation between the replays
            other.ChannelIterate(1)
            b.ChannelIterate(1)
            for f in fields:
                assert np.array_equal(view.download(f), b.download(f)), (kind, step, f)
            assert view.coe()[7] == pytest.approx(b.coe()[7], rel=0, abs=1e-14)  # SimTime
        p50, p99, n = ens.latency()
        assert n == 4 and 0.0 < p50 <= p99
        got = ens.release(cid, keep=True)
        got.ChannelIterate(2)  # a released context steps on its own again (its own step graph)
        b.ChannelIterate(2)
        assert np.array_equal(got.download(capi.F_UX), b.download(capi.F_UX))
        got.close()
    other.close()
    b.close()


def test_admission_follows_the_hbm_budget():
    case = CS.case_channel(32, 24, 16, 1.8)
    s1, s2 = twins(case)
    fp = capi.load()
    import ctypes as C
    need = C.c_size_t()
    capi.check(fp.ptc_case_footprint(s1.h, C.byref(need)), "ptc_case_footprint")
    assert need.value >= 10 * 8 * 32 * 26 * 16  # at least the ten fields a step touches
    with E.Ensemble(int(1.5 * need.value)) as ens:
        c1 = ens.admit(s1)
        assert c1 is not None
        assert ens.admit(s2) is None  # over budget: the caller keeps the context
        used, budget, live = ens.usage()
        assert live == 1 and used == need.value
        ens.step(2)
        assert ens.steps_done(c1) == 2
        ens.release(c1)
        c2 = ens.admit(s2)
        assert c2 is not None
        ens.step(1)
        p50, p99, n = ens.latency(reset=True)
        assert n == 3
        assert ens.latency()[2] == 0


def test_many_cases_overlap_and_finish():
    wl = E.c3_workload(ncases=6)
    solvers = [E.make_case(c, 10 + q) for q, c in enumerate(wl.cases)]
    ref = [E.make_case(c, 10 + q) for q, c in enumerate(wl.cases)]
    with E.Ensemble(1 << 36) as ens:
        ids = [ens.admit(s) for s in solvers]
        assert None not in ids
        ens.step(3)
        for q, cid in enumerate(ids):
            got = ens.release(cid, keep=True)
            ref[q].ChannelIterate(3)
            assert np.array_equal(got.download(capi.F_UX), ref[q].download(capi.F_UX)), q
            got.close()
            ref[q].close()


def test_ensemble_graph_recaptured_after_buffer_changes():
    """An admitted case's captured step holds the context's reduction
    partials and Poisson plan buffers.  clcStat on the admitted context
    (growing the partials) and a Poisson plan rebuild (freeing the plan's
    buffers) between ensemble steps make the next step re-capture; the
    fields stay those of ordinary launches, bit for bit."""
    case = CS.case_channel(64, 48, 64, 1.8, re_b=3000.0)
    a, b = twins(case, seed=7)
    b.step_graphs(False)
    with E.Ensemble(1 << 34) as ens:
        cid = ens.admit(a)
        view = ens.view(cid)
        ens.step(1)
        b.ChannelIterate(1)
        view.clcStat()
        b.clcStat()
        ens.step(2)
        b.ChannelIterate(2)
        for f in FIELDS:
            assert np.array_equal(view.download(f), b.download(f)), f
        view.poisson_config(capi.PREC_FP64, capi.TRI_AUTO, 0, 0)  # frees and rebuilds the plan
        b.poisson_config(capi.PREC_FP64, capi.TRI_AUTO, 0, 0)
        ens.step(1)
        b.ChannelIterate(1)
        for f in FIELDS:
            assert np.array_equal(view.download(f), b.download(f)), f
        assert np.array_equal(view.get_stat()[0], b.get_stat()[0])
    b.close()


@pytest.mark.parametrize("nsteps,tol", [(1, 1e-12), (5, 1e-10)])
def test_c3_draws_against_the_oracle(nsteps, tol):
    """The first four draws of the 64-case workload (seed 3: 64x96x128,
    64x128x96 with the settling scalar, 128x64x128, 64x96x192 with the
    settling scalar), admitted together and stepped through the ensemble,
    against the oracle's ChannelIterate on the same initial fields (noise
    plus one projection, a uniform scalar of 1)."""
    from helpers import inner_y, oracle_project, rel_err
    from oracle.oracle import Oracle, State
    from paratc_amd import ic

    wl = E.c3_workload(ncases=4)
    assert sum(c.scalar for c in wl.cases) == 2
    seeds = [100 + q for q in range(4)]
    solvers = [E.make_case(c, sd) for c, sd in zip(wl.cases, seeds)]
    refs = []
    for case, sd in zip(wl.cases, seeds):
        o = Oracle(case)
        st = State(o)
        st.ux, st.uy, st.uz = ic.channel_fields(case, seed=sd, amp=0.1)
        o.SetBC_and_UpdateHalo(st.ux, st.uy, st.uz, st.uy_ym)
        if case.scalar:
            st.scalar[:] = 1.0
            o.SetBC_and_UpdateHalo_scalar(st.scalar)
        assert oracle_project(o, st) <= 1e-12
        refs.append((o, st))
    with E.Ensemble(1 << 36) as ens:
        ids = [ens.admit(s) for s in solvers]
        assert None not in ids
        ens.step(nsteps)
        for q, cid in enumerate(ids):
            o, st = refs[q]
            for _ in range(nsteps):
                o.ChannelIterate(st, 1)
            view = ens.view(cid)
            pairs = [(capi.F_UX, st.ux), (capi.F_UY, st.uy), (capi.F_UZ, st.uz), (capi.F_PR, st.p)]
            if wl.cases[q].scalar:
                pairs.append((capi.F_SCALAR, st.scalar))
            for fid, ref in pairs:
                e = rel_err(inner_y(view.download(fid)), inner_y(ref))
                # p: the ill-conditioned low modes, as in the single-case tests
                bar = tol if fid != capi.F_PR else 10 * tol
                print(f"C3 draw {q} ({wl.cases[q].nx}x{wl.cases[q].ny}x{wl.cases[q].nz}"
                      f"{', scalar' if wl.cases[q].scalar else ''}) {nsteps} steps field {fid}: {e:.3e}",
                      flush=True)
                assert e <= bar, (q, fid, e)
            assert view.CheckDivergence(capi.F_UX, capi.F_UY, capi.F_UZ) <= 1e-10
