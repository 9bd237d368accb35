"""GPU checks of the seeded input generators (ptc_seed_normal,
ptc_seed_fields) against the host generator (paratc_amd/rng.py, its
SplitMix64 counter stream pinned by known answers in tests/test_ic.py):
uniform-based fields bit for bit (the tanh front to the last bits of the
device tanh), normals to the last bits of the device log, sqrt, cos and
sin."""
import numpy as np
import pytest

from helpers import rel_err
from paratc_amd import capi, case as CS, ic
from paratc_amd.solver import Solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("grid", [(32, 24, 16), (36, 10, 14)], ids=lambda g: "x".join(map(str, g)))
@pytest.mark.parametrize("compatible", [False, True])
def test_seed_normal_matches_host(grid, compatible):
    case = CS.case_poisson(*grid, gamma=2.0)
    with Solver(case) as s:
        s.seed_normal(capi.F_ARR1, seed=4, stream=1, compatible=compatible)
        got = s.download(capi.F_ARR1, capi.L_INTERIOR)
    if compatible:
        ref = ic.poisson_rhs(case, seed=4, stream=1)
    else:
        from paratc_amd import rng
        ref = rng.normal(4, case.npts, 1).reshape((case.nx, case.ny, case.nz), order="F")
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    if compatible:
        dyp = np.diff(case.yp())
        assert abs((got.sum(axis=(0, 2)) * dyp).sum()) <= 1e-12 * got.size


def test_seed_channel_fields_match_host():
    case = CS.case_channel(48, 24, 16, 1.8)
    with Solver(case) as s:
        s.seed_fields(capi.SEED_CHANNEL, seed=7, amp=0.1)
        got = [s.download(f) for f in (capi.F_UX, capi.F_UY, capi.F_UZ)]
    ref = ic.channel_fields(case, seed=7, amp=0.1)
    for g, r in zip(got, ref):
        # the device layout stores no x/z halo: the download fills the images
        assert np.array_equal(g[2:-2, :, 2:-2], r[2:-2, :, 2:-2])


def test_seed_front_fields_match_host():
    case = CS.case_gravity_current(64, 24, 16, gamma=1.6)
    with Solver(case) as s:
        s.seed_fields(capi.SEED_FRONT, seed=2, amp=1e-3)
        got = [s.download(f) for f in (capi.F_UX, capi.F_UY, capi.F_UZ, capi.F_SCALAR)]
    ref = ic.gravity_current_fields(case, seed=2, amp=1e-3)
    for g, r in zip(got[:3], ref[:3]):
        assert np.array_equal(g[2:-2, :, 2:-2], r[2:-2, :, 2:-2])
    assert rel_err(got[3][2:-2, :, 2:-2], ref[3][2:-2, :, 2:-2]) <= 1e-15
