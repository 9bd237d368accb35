"""GPU parity of the dump_visu fields (src/m_IOAndVisu.f90:182-385) with the
oracle: cell-centre velocities (ux with uCRF), pressure, scalar, vorticity
components and magnitude, the vortex criteria Q and lambda2, with and without
subsampling, FP64 and FP32."""
import numpy as np
import pytest

from helpers import rel_err
from oracle.oracle import Oracle
from paratc_amd import capi
from paratc_amd.solver import Solver
from test_gpu_tscheme import rand_state, small_channel, small_gravity, upload_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["channel", "gravity"])
@pytest.mark.parametrize("skip", [(1, 1, 1), (2, 3, 2)], ids=["all", "skip232"])
def test_visu_fields_match_oracle(kind, skip):
    case = small_channel(40, 24, 18) if kind == "channel" else small_gravity(48, 24, 16)
    orc = Oracle(case)
    ux, uy, uz, p, sc = rand_state(case, orc, seed=21)
    nf = 9 if case.scalar else 8
    with Solver(case) as s:
        upload_state(s, ux, uy, uz, p, sc)
        for which in list(range(nf)) + [capi.VISU_Q, capi.VISU_LAMBDA2]:
            ref = orc.dump_visu(which, ux, uy, uz, p, sc, skip=skip)
            got = s.visu_field(which, skip, 8)
            assert got.shape == ref.shape
            e = rel_err(got, ref)
            print(kind, skip, which, f"{e:.2e}", flush=True)
            # lambda2 goes through acos, cos and cube roots: device and host
            # libm differ by ulps, amplified where two eigenvalues nearly meet
            assert e <= (1e-10 if which == capi.VISU_LAMBDA2 else 1e-13), which
            got4 = s.visu_field(which, skip, 4)
            assert got4.dtype == np.float32
            assert np.allclose(got4, ref.astype(np.float32), rtol=2e-7, atol=1e-7 * np.abs(ref).max()), which
        if not case.scalar:
            with pytest.raises(capi.PtcError, match="scalar"):
                s.visu_field(capi.VISU_SCALAR, skip, 8)
        out = np.zeros(5)
        rc = s.L.ptc_visu_field(s.h, capi.VISU_UX, 1, 1, 1, 8, out.ctypes.data, out.size)
        assert rc != 0 and b"values" in s.L.ptc_last_error()
