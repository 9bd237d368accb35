"""GPU parity of the clcStat accumulation (src/m_FlowCase.f90:283-428)."""
import numpy as np
import pytest

from helpers import random_halo
from oracle.oracle import Oracle
from paratc_amd import capi, case as CS
from paratc_amd import ensemble as E
from paratc_amd.solver import Solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["channel", "scalar"])
def test_stat_accumulation_matches_oracle(kind):
    if kind == "scalar":
        case = E.case_channel_scalar(32, 24, 16, re_b=3000.0, ws=0.02)
    else:
        case = CS.case_channel(48, 20, 24, 1.7)
        case.is_use_crf, case.ucrf_in = True, 0.2
    orc = Oracle(case)
    ns = 40 if case.scalar else 35
    ref = np.zeros((ns, case.ny + 1), order="F")
    with Solver(case) as s:
        for sample in range(2):
            # x/z halos hold the periodic images, as the device layout implies
            f = [random_halo((case.nx, case.ny, case.nz), 30 + 5 * sample + q) for q in range(5)]
            for fid, a in zip((capi.F_UX, capi.F_UY, capi.F_UZ, capi.F_PR, capi.F_SCALAR), f):
                if fid != capi.F_SCALAR or case.scalar:
                    s.upload(fid, a)
            orc.clcStat(*f, ref)
            s.clcStat()
        got, nf, pgsum = s.get_stat(reset=True)
        assert nf == 2
        again, nf0, pg0 = s.get_stat()
    scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-300)
    err = np.abs(got - ref) / scale
    print(kind, "max err per stat", float(err.max()), flush=True)
    assert err.max() <= 1e-12
    assert nf0 == 0 and not again.any() and pg0 == 0.0
