"""The allocation-free clcPrSrc (ptc_clc_prsrc_noalloc) against the ordinary
entry point ptc_clc_prsrc, both called through the C-ABI: the same Poisson
source bit for bit and the same divergence maximum, on two calls in a row (the
second on a smaller grid, so the shared device scratch holds the first call's
larger partials), then the limit and state errors."""
import ctypes as C

import numpy as np
import pytest

from helpers import random_halo
from paratc_amd import capi, case as CS
from paratc_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _both(case, seed):
    """Velocities with random interiors and wall rows; the source and divmax
    of each entry point, the ordinary one into RealArr1 and the
    allocation-free one into the spare work array."""
    shape = (case.nx, case.ny, case.nz)
    with Solver(case) as s:
        for q, fid in enumerate((capi.F_UX, capi.F_UY, capi.F_UZ)):
            s.upload(fid, random_halo(shape, seed, amp=1.0, stream=q))
        s.PMcoeUpdate(2, case.dtmax)
        s.fill(capi.F_ARR1, 0.0)
        s.fill(capi.F_WORK1, 7.0)
        ref = C.c_double()
        got = C.c_double()
        capi.check(s.L.ptc_clc_prsrc(s.h, capi.F_UX, capi.F_UY, capi.F_UZ, capi.F_ARR1, C.byref(ref)),
                   "clcPrSrc")
        capi.check(s.L.ptc_clc_prsrc_noalloc(s.h, capi.F_UX, capi.F_UY, capi.F_UZ, capi.F_WORK1,
                                             C.byref(got)), "clcPrSrc without allocation")
        f_ref = s.download(capi.F_ARR1, capi.L_INTERIOR)
        f_got = s.download(capi.F_WORK1, capi.L_INTERIOR)
    return f_ref, f_got, ref.value, got.value


def test_noalloc_prsrc_matches_prsrc_twice():
    # first call on the larger grid, then a smaller one
    for grid, seed in (((96, 64, 48), 11), ((32, 24, 16), 12)):
        case = CS.case_channel(*grid, 1.8)
        f_ref, f_got, d_ref, d_got = _both(case, seed)
        print(f"{grid}: divmax {d_ref:.6e} / {d_got:.6e}", flush=True)
        assert np.array_equal(f_got, f_ref)
        assert d_got == d_ref and d_ref > 0.0


def test_noalloc_prsrc_limits():
    L = capi.load()
    # 9216 x 140 x 1400: 144 * 35 * 1400 partials, above the 2^21 the scratch holds
    big = CS.case_poisson(9216, 140, 1400, 2.2)
    with Solver(big) as s:
        rc = L.ptc_clc_prsrc_noalloc(s.h, capi.F_UX, capi.F_UY, capi.F_UZ, capi.F_ARR1, None)
        assert rc == capi.PTC_E_ARG
        assert b"too large" in L.ptc_last_error()
    small = CS.case_channel(32, 24, 16, 1.8)
    with Solver(small) as s:
        # the fields have no device memory yet: no allocation, a state error
        rc = L.ptc_clc_prsrc_noalloc(s.h, capi.F_UX, capi.F_UY, capi.F_UZ, capi.F_ARR1, None)
        assert rc == capi.PTC_E_STATE
        assert L.ptc_clc_prsrc_noalloc(None, 0, 1, 2, 9, None) == capi.PTC_E_ARG

