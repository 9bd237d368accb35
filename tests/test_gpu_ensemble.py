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
