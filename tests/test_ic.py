"""InitVelocity (src/m_FlowCase.f90:49-107) restated in paratc_amd.ic with a
seeded Deviation: its invariants fixed by the reference's construction."""
import numpy as np

from oracle import oracle as O
from paratc_amd import case as CS, ic


def test_init_velocity_invariants():
    """After InitVelocity the oracle's CalcUxAver of u is ubulk - uCRF (the
    ratiot scaling, then the shift), every x-z plane mean of w is zero, v is
    zero, the streak amplitude lies within the Deviation band, and the seed
    fixes the field."""
    for crf in (False, True):
        case = CS.case_channel(32, 24, 16, 1.9)
        case.is_use_crf, case.ucrf_in = crf, 0.3
        o = O.Oracle(case)
        ux, uy, uz = ic.init_velocity(case, seed=4)
        assert ux.shape == (case.nx + 4, case.ny + 2, case.nz + 4)
        assert abs(o.CalcUxAver(ux) - (case.ubulk - case.ucrf)) <= 1e-14
        core = (slice(2, -2), slice(1, -1), slice(2, -2))
        assert np.abs(uz[core].mean(axis=(0, 2))).max() <= 1e-15
        assert not uy.any()
        # ghost rows stay zero until the boundary routines fill them
        assert not ux[:, 0, :].any() and not ux[:, -1, :].any()
        # periodic x and z halos
        assert np.array_equal(ux[0], ux[case.nx]) and np.array_equal(uz[:, :, 1], uz[:, :, case.nz + 1])
        ux2, _, uz2 = ic.init_velocity(case, seed=4)
        assert np.array_equal(ux, ux2) and np.array_equal(uz, uz2)
        ux3, _, _ = ic.init_velocity(case, seed=5)
        assert not np.array_equal(ux, ux3)


def test_init_velocity_streak_wavelength_fits_the_box():
    """w is A(y) sin(wx x+) Dev minus its plane mean, with wx rounded so that a
    whole number of wavelengths fits the box: along x its sign changes an even
    number of times, counting the periodic wrap."""
    case = CS.case_channel(64, 16, 8, 1.5)
    _, _, uz = ic.init_velocity(case, seed=1)
    w = uz[2:-2, 1:-1, 2:-2]
    s = np.sign(w[:, case.ny // 2, 0])
    flips = int((s[1:] != s[:-1]).sum() + (s[0] != s[-1]))
    assert flips % 2 == 0 and flips >= 2
