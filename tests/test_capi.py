"""CPU checks of the C-ABI boundary and the host-side helpers (no GPU compute)."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from paratc_amd import capi, case as CS, ic, rng

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "paratc_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ptc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = capi.load()
    syms = header_symbols()
    assert len(syms) >= 40
    assert sorted(capi.EXPORTS) == syms
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (ptc_[a-z0-9_]+)", out))
    assert set(syms) <= exported


def test_abi_version_and_error_text():
    L = capi.load()
    assert L.ptc_abi_version() == 1
    assert isinstance(L.ptc_last_error(), bytes)


def _create(case_mod):
    L = capi.load()
    g, p = capi.params_from_case(CS.case_channel(32, 24, 16))
    case_mod(g, p)
    yp = np.ascontiguousarray(CS.case_channel(32, 24, 16).yp())
    h = C.c_void_p()
    rc = L.ptc_create(C.byref(g), C.byref(p), yp.ctypes.data_as(C.POINTER(C.c_double)), C.byref(h))
    return rc, L.ptc_last_error().decode(), h


@pytest.mark.parametrize("mod,msg", [
    (lambda g, p: setattr(g, "nx", 33), "mod(nxc,2)"),
    (lambda g, p: setattr(g, "ny", 23), "mod(nyc,2)"),
    (lambda g, p: setattr(p, "ischeme", 4), "Time scheme"),
    (lambda g, p: setattr(g, "nx", 44), "factor into 2, 3, 5 and 7"),
    (lambda g, p: setattr(p, "flow_type", 3), "FlowType"),
])
def test_create_rejects_bad_input_before_touching_a_device(mod, msg):
    rc, err, h = _create(mod)
    assert rc == capi.PTC_E_ARG or rc == capi.PTC_E_UNSUPPORTED
    assert msg in err
    assert not h.value


def test_null_context_is_an_error():
    L = capi.load()
    assert L.ptc_clc_ppe(None, capi.F_ARR1, capi.F_HALO) == capi.PTC_E_ARG
    assert b"null context" in L.ptc_last_error()


def test_splitmix64_known_answers():
    # SplitMix64 from seed 0 (Steele, Lea, Flood 2014; Vigna's reference outputs)
    assert [int(v) for v in rng.splitmix64(0, 3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    # counter form: any window equals the sequence
    assert np.array_equal(rng.splitmix64(5, 4, start=6), rng.splitmix64(5, 10)[6:])
    z = rng.normal(1, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def test_poisson_rhs_is_compatible_and_seeded():
    case = CS.case_poisson(32, 24, 16, 2.0)
    f = ic.poisson_rhs(case, seed=1)
    yp = case.yp()
    dyp = np.diff(yp)
    assert abs((f.sum(axis=(0, 2)) * dyp).sum()) < 1e-12 * np.abs(f).sum()
    assert np.array_equal(f, ic.poisson_rhs(case, seed=1))
    assert not np.array_equal(f, ic.poisson_rhs(case, seed=1, stream=1))


def test_channel_fields_profile():
    case = CS.case_channel(32, 24, 16, 1.8)
    ux, uy, uz = ic.channel_fields(case, seed=0, amp=0.0)
    yc = 0.5 * (case.yp()[1:] + case.yp()[:-1])
    assert np.allclose(ux[5, 1:-1, 5], 1.5 * (1 - (yc - 1) ** 2))
    assert not uy.any() and not uz.any()
    ux, uy, uz = ic.channel_fields(case, seed=0, amp=0.1)
    assert np.all(uy[:, 1, :] == 0.0) and np.abs(uz).max() <= 0.1


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_restart_roundtrip_and_size_check(tmp_path):
    from paratc_amd import restart

    r = np.random.default_rng(0)
    f = [r.standard_normal((6, 4, 8)) for _ in range(5)]
    path = str(tmp_path / restart.restart_name("", "Cha", 12).lstrip("/"))
    assert path.endswith("RestartForCha0000000012")
    restart.write_restart(path, *f[:4], scalar=f[4], ucrf=0.25)
    back = restart.read_restart(path, 6, 4, 8, scalar=True, ucrf=0.25)
    assert np.allclose(back["ux"], f[0]) and np.array_equal(back["scalar"], f[4])
    assert os.path.getsize(path) == 8 * 6 * 4 * 8 * 5
    with open(path, "ab") as fh:  # truncated / padded file -> "file byte wrong"
        fh.write(b"\0" * 8)
    with pytest.raises(ValueError, match="file byte wrong2"):
        restart.read_restart(path, 6, 4, 8, scalar=True)


def test_ensemble_workload_draw():
    from paratc_amd import ensemble as E

    wl = E.c3_workload()
    assert len(wl.cases) == 64 and sum(c.scalar for c in wl.cases) == 32
    assert all(c.nx in E.NX_SET and c.ny in E.NY_SET and c.nz in E.NZ_SET for c in wl.cases)
    assert all(1 / 5000.0 <= c.xnu <= 1 / 2000.0 for c in wl.cases)
    assert all(0.005 <= c.falling_vel <= 0.03 for c in wl.cases if c.scalar)
    gaps = np.diff(wl.arrivals)
    assert np.all(gaps >= 0) and 0.5 / 20 < gaps.mean() < 2.0 / 20  # Poisson process, rate 20/s
    again = E.c3_workload()
    assert [(c.nx, c.ny, c.nz, c.xnu) for c in again.cases] == [(c.nx, c.ny, c.nz, c.xnu) for c in wl.cases]


def test_fortran_binding_names_declared_symbols():
    text = open(os.path.join(ROOT, "paratc_amd", "fortran", "m_paratc_gpu.f90")).read()
    names = set(re.findall(r'bind\(C,\s*name="(ptc_[a-z0-9_]+)"\)', text))
    assert len(names) >= 30
    assert names <= set(header_symbols()), names - set(header_symbols())


def test_stats_file_format(tmp_path):
    from paratc_amd import stats

    st = np.arange(35 * 3, dtype=float).reshape((35, 3), order="F") - 7.5
    path = str(tmp_path / "s")
    stats.write_stats(path, 400, 4, 100, True, -0.0123, st)
    lines = open(path).read().splitlines()
    assert lines[0] == "  The time step range for this fluid statistics is     100:    100:    400"
    assert lines[3] == "    time averaged pressure gradient is:    3.075000000000000E-03"
    assert len(lines) == 5 + 3 and all(len(ln) == 24 * 35 for ln in lines[5:])
    assert float(lines[5][:24]) == -7.5 / 4 and stats.stats_name("r/", 7) == "r/stats0000000007"
