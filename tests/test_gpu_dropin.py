"""The drop-in Fortran modules (paratc_amd/fortran/dropin: m_Poisson,
m_TScheme, m_BC_and_Halo, m_Tools under the reference's names and argument
lists) driven by a Fortran host that runs ChannelInitialize's and
ChannelIterate's calls (src/m_ChannelSystem.f90:41-66, 80-145) on its own
arrays (tests/fortran/dropin_iterate.f90), plain and -DScalarFlow builds;
every field after several steps, SimTime and PrGradAver against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from helpers import inner, inner_y, rel_err
from oracle.oracle import Oracle, State
from paratc_amd import case as CS, ic

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = {False: os.path.join(ROOT, "tests", "fortran", "dropin_channel"),
       True: os.path.join(ROOT, "tests", "fortran", "dropin_scalar")}


def _nml_list(v):
    return ", ".join(repr(float(x)) for x in v)


def _run_host(tmp_path, case, nsteps, fields):
    exe = EXE[bool(case.scalar)]
    assert os.path.exists(exe), "drop-in test host not built (run __graft_entry__.build())"
    case.yp().astype(np.float64).tofile(tmp_path / "yp.bin")
    with open(tmp_path / "fields.in", "wb") as fh:
        for a in fields:
            fh.write(np.asfortranarray(inner(a)).tobytes(order="F"))
    xb, yb, zb = case.bc_values()
    gev, fvv = case.scalar_vectors()
    (tmp_path / "run.nml").write_text(
        "&dropin\n"
        f" nx={case.nx}, ny={case.ny}, nz={case.nz}, xlx={case.xlx!r}, yly={case.yly!r}, zlz={case.zlz!r},\n"
        f" xnu={case.xnu!r}, ubulk={case.ubulk!r}, ucrf={case.ucrf!r}, dtmax={case.dtmax!r}, nsteps={nsteps},\n"
        f" flow_type={case.flow_type}, is_ux_const={int(case.is_ux_const)}, ischeme={case.ischeme},\n"
        f" icfl_in={case.icfl}, cflc_in={case.cflc!r},\n"
        f" uxbc={_nml_list(xb)}, uybc={_nml_list(yb)}, uzbc={_nml_list(zb)}, grav={_nml_list(case.gravity)},\n"
        f" sdiff={case.sdiff!r}, gev={_nml_list(gev)}, fvv={_nml_list(fvv)},\n"
        f" sbc_opt={case.scalar_bc_option[0]}, {case.scalar_bc_option[1]}, sbc_val={_nml_list(case.scalar_bc_values)},\n"
        f" yp_file='{tmp_path / 'yp.bin'}', fields_in='{tmp_path / 'fields.in'}',\n"
        f" fields_out='{tmp_path / 'fields.out'}'\n/\n")
    r = subprocess.run([exe, str(tmp_path / "run.nml")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    # InitPoissonSolver prints its search, as the reference prints its choices' times
    assert "Auto-tuning mode for Poisson Solver" in r.stdout and "  picked " in r.stdout, r.stdout
    nf = 5 if case.scalar else 4
    shape = (case.nx, case.ny + 2, case.nz)
    raw = np.fromfile(tmp_path / "fields.out")
    n = int(np.prod(shape))
    out = [raw[q * n:(q + 1) * n].reshape(shape, order="F") for q in range(nf)]
    simtime, prgrad = raw[nf * n], raw[nf * n + 1]
    mon = raw[nf * n + 2:nf * n + 9]
    return out, simtime, prgrad, mon


@pytest.mark.parametrize("kind,nsteps", [("channel", 3), ("gravity", 2)])
def test_dropin_host_matches_oracle(tmp_path, kind, nsteps):
    if kind == "channel":
        case = CS.case_channel(32, 24, 16, 1.8)
        ux, uy, uz = ic.channel_fields(case, seed=0, amp=0.1)
        sc = None
    else:
        case = CS.case_gravity_current(48, 24, 16, gamma=1.6)
        case.scalar_bc_option = (-2, -3)
        case.scalar_bc_values = (3.5, 0.0, 0.2, 0.0)
        case.is_ux_const, case.ubulk = True, 0.5
        case.ucrf_in, case.is_use_crf = 0.2, True
        ux, uy, uz, sc = ic.gravity_current_fields(case, seed=2, amp=1e-2)
    p0 = np.zeros_like(ux)
    fields = [ux, uy, uz, p0] + ([sc] if sc is not None else [])
    out, simtime, prgrad, mon = _run_host(tmp_path, case, nsteps, fields)

    orc = Oracle(case)
    st = State(orc)
    st.ux, st.uy, st.uz = ux.copy(order="F"), uy.copy(order="F"), uz.copy(order="F")
    if sc is not None:
        st.scalar = sc.copy(order="F")
    orc.SetBC_and_UpdateHalo(st.ux, st.uy, st.uz, st.uy_ym)
    orc.SetBC_and_UpdateHalo_pr(st.p)
    if sc is not None:
        orc.SetBC_and_UpdateHalo_scalar(st.scalar)
    for _ in range(nsteps):
        mon_o = orc.ChannelIterate(st, 1)
    coe = orc.coe()
    refs = [st.ux, st.uy, st.uz, st.p] + ([st.scalar] if sc is not None else [])
    names = ["ux", "uy", "uz", "p", "scalar"]
    for name, got, ref in zip(names, out, refs):
        e = rel_err(got, inner_y(ref))
        print(f"{kind} {nsteps} steps {name}: {e:.3e}", flush=True)
        # p carries the ill-conditioned low modes of the solve (see the step tests)
        assert e <= (1e-10 if name == "p" else 1e-12), name
    assert abs(simtime - coe[7]) <= 1e-14 * max(1.0, abs(coe[7]))
    assert abs(prgrad - coe[8]) <= 1e-10 * max(1.0, abs(coe[8])), (prgrad, coe[8])
    assert mon[1] <= 1e-10 and abs(mon[1] - mon_o[1]) <= 1e-12
