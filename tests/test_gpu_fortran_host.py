"""The Fortran host (paratc_amd/fortran/ptc_channel, built with amdflang)
drives the GPU through the iso_c_binding module, reading and writing the
reference's restart format; its result must match the oracle."""
import os
import subprocess

import numpy as np
import pytest

from helpers import inner, rel_err
from oracle.oracle import Oracle, State
from paratc_amd import case as CS, ic, restart

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "paratc_amd", "fortran", "ptc_channel")


def test_fortran_host_matches_oracle(tmp_path):
    assert os.path.exists(EXE), "Fortran host not built (run __graft_entry__.build())"
    case = CS.case_channel(32, 24, 16, 1.8)
    nsteps = 2
    ux, uy, uz = ic.channel_fields(case, seed=0, amp=0.1)
    p0 = np.zeros((case.nx, case.ny, case.nz), order="F")
    rin, rout = tmp_path / "RestartForGPU0000000000", tmp_path / "RestartForGPU0000000002"
    restart.write_restart(str(rin), inner(ux), inner(uy), inner(uz), p0)
    case.yp().astype(np.float64).tofile(tmp_path / "yp.bin")
    (tmp_path / "run.nml").write_text(
        "&ptc_run\n"
        f" nx={case.nx}, ny={case.ny}, nz={case.nz}, xlx={case.xlx!r}, yly={case.yly!r}, zlz={case.zlz!r},\n"
        f" xnu={case.xnu!r}, ubulk={case.ubulk!r}, dtmax={case.dtmax!r}, nsteps={nsteps},\n"
        f" flow_type=1, is_ux_const=1, ischeme=3,\n"
        f" yp_file='{tmp_path / 'yp.bin'}', restart_in='{rin}', restart_out='{rout}',\n"
        f" spec_out='{tmp_path / 'final_'}', jforlcs=3\n/\n")
    r = subprocess.run([EXE, str(tmp_path / "run.nml")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    got = restart.read_restart(str(rout), case.nx, case.ny, case.nz)

    orc = Oracle(case)
    st = State(orc)
    st.ux, st.uy, st.uz = ux, uy, uz
    orc.SetBC_and_UpdateHalo(st.ux, st.uy, st.uz, st.uy_ym)
    orc.SetBC_and_UpdateHalo_pr(st.p)
    for _ in range(nsteps):
        orc.ChannelIterate(st, 1)
    for name, ref in (("ux", st.ux), ("uy", st.uy), ("uz", st.uz)):
        assert rel_err(got[name], inner(ref)) <= 1e-12, name
    assert rel_err(got["p"], inner(st.p)) <= 1e-10
    # clcStat's 1-D spectra of the final fields, written by the host
    sx, sz = orc.spectra_1d(st.ux, st.uy, st.uz, st.p, jlcs=3)
    fx = np.fromfile(tmp_path / "final_SpecX").reshape(sx.shape, order="F")
    fz = np.fromfile(tmp_path / "final_SpecZ").reshape(sz.shape, order="F")
    for m in range(8):
        assert rel_err(fx[..., m], sx[..., m]) <= 1e-9, m
        assert rel_err(fz[..., m], sz[..., m]) <= 1e-9, m
