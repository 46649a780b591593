"""CPU checks of the C-ABI library and the host-side helpers (no compute calls on a GPU)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "cntc_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"\bvoid\s+(cntc_\w+)\s*\(", text)


def test_library_exports_every_header_symbol(cntc):
    names = header_functions()
    assert len(names) >= 30
    lib = ctypes.CDLL(cntc.library_path)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror declares a prototype for each of them
    assert sorted(set(names) - set(cntc._PROTOS)) == []


def test_exported_symbols_are_plain_c(cntc):
    out = subprocess.run(["nm", "-D", "--defined-only", cntc.library_path], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    for n in header_functions():
        assert n in exported, n


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_calculate_without_gpu_fails_loudly(cntc):
    """No CPU fallback: without a GPU the solve reports CNTC_err_other."""
    from contact_amd import workloads as W
    c = W.hertz_40()
    W.apply_case(cntc, 7, 1, c)
    assert cntc.calculate(7, 1) == -99
    err = cntc.calculatebatch([7], [1])
    assert list(err) == [-99]


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_workloads_are_seeded():
    from contact_amd import workloads as W
    a = W.hertz_batch(n=2048)
    b = W.hertz_batch(n=2048)
    assert [(c["mx"], c["my"], c["fn"]) for c in a] == [(c["mx"], c["my"], c["fn"]) for c in b]
    sides = np.array([[c["mx"], c["my"]] for c in a])
    assert sides.min() >= 20 and sides.max() <= 256
    ratio = np.array([c["aa"] / c["bb"] for c in a])
    assert ratio.min() >= 0.2 - 1e-12 and ratio.max() <= 5.0 + 1e-12
    fn = np.array([c["fn"] for c in a])
    assert fn.min() >= 2e4 and fn.max() <= 1.5e5
    r1 = W.rough_patch(n=64)
    r2 = W.rough_patch(n=64)
    assert np.array_equal(r1["h"], r2["h"])


def test_roughness_statistics():
    from contact_amd import workloads as W
    f = W.roughness(256, 256, 1e-3, 4.0, 1234)
    assert abs(f.std() - 1e-3) < 1e-12 and abs(f.mean()) < 1e-12
    # exponential correlation: correlation at lag 4 elements near exp(-1)
    c4 = np.mean(f * np.roll(f, 4, axis=1)) / f.var()
    assert 0.2 < c4 < 0.5


def test_hertz_semiaxes_matches_oracle(oracle):
    """The workload generator's Hertz solution against the oracle's (hzcalc3d)."""
    from contact_amd import workloads as W
    from tests import cases as C
    a1, b1 = 1.0 / 920.0, 1.0 / 600.0
    a, b = W.hertz_semiaxes(a1, b1, 1e5)
    P = oracle.Problem()
    c = C.hertz_case(a, b, 20, 20, 1e5)
    P.calculate(C.to_oracle(oracle, c)[0])
    oa1, ob1 = P.hertz()[:2]
    assert oa1 == pytest.approx(a1, rel=1e-7) and ob1 == pytest.approx(b1, rel=1e-7)


def test_flags_parameters_and_velocity_round_trip(cntc):
    """cntc_getflags / cntc_getparameters / cntc_getreferencevelocity return what the set-calls
    stored (src/contact_addon.f90:4001-4258, 4987-5017), with ic_init's defaults for digits
    never set (src/m_hierarch_data.f90:966-1014)."""
    K = cntc.CNTC
    ire, icp = 41, 2
    cntc.setflags(ire, icp, [K["ic_tang"], K["ic_norm"], K["ic_pvtime"], K["ic_iestim"],
                             K["ic_force"], K["ic_matfil"], K["ic_flow"], K["ic_exrhs"]],
                  [2, 1, 0, 3, 1, 2, 7, 3])
    got = cntc.getflags(ire, icp, [K["ic_tang"], K["ic_norm"], K["ic_pvtime"], K["ic_iestim"],
                                   K["ic_force"], K["ic_matfil"], K["ic_flow"], K["ic_config"],
                                   K["ic_bound"], K["ic_output"], K["ic_return"], 0])
    assert list(got) == [2, 1, 0, 3, 1, 2, 7, 1, 0, 2, 1, 0]
    cntc.setmaterialparameters(ire, icp, 0, [0.28, 0.30, 82000.0, 70000.0])
    m = cntc.getparameters(ire, icp, 2, 7)
    ga = 2.0 / (1.0 / 82000.0 + 1.0 / 70000.0)
    nu = ga * (0.28 / 82000.0 + 0.30 / 70000.0) / 2.0
    ak = ga / 4.0 * ((1 - 0.56) / 82000.0 - (1 - 0.60) / 70000.0)
    assert m == pytest.approx([82000.0, 70000.0, ga, 0.28, 0.30, nu, ak], rel=1e-14)
    cntc.setfrictionmethod(ire, icp, 0, [0.35, 0.25])
    f = cntc.getparameters(ire, icp, 3, 7)
    assert f[1] == 1.0 and f[5] == 0.35 and f[6] == 0.25
    cntc.setreferencevelocity(ire, icp, 12345.0)
    assert cntc.getreferencevelocity(ire, icp) == 12345.0
    # SI units: velocity in m/s, moduli in Pa
    cntc.setflags(ire, 3, [K["if_units"]], [K["un_si"]])
    cntc.setreferencevelocity(ire, 3, 30.0)
    assert cntc.getreferencevelocity(ire, 3) == pytest.approx(30.0, rel=1e-15)
    cntc.setmaterialparameters(ire, 3, 0, [0.28, 0.28, 8.2e10, 8.2e10])
    assert cntc.getparameters(ire, 3, 2, 3) == pytest.approx([8.2e10] * 3, rel=1e-15)
    # metadata are accepted (and change nothing the getters report)
    cntc.setmetadata(ire, icp, [K["mt_tim"], 0, K["mt_run"], 999], [1.5, 0.0, 3.0, 1.0])
    assert cntc.getreferencevelocity(ire, icp) == 12345.0
    cntc.finalize(ire)


def test_hertz_contact_options_checked_in_reference_order(cntc):
    """cntc_sethertzcontact checks the method before the parameter count
    (src/contact_addon.f90:2003-2014): -6 takes 6 parameters, -1..-5 take 5, others none."""
    ire, icp = 42, 1
    cntc.sethertzcontact(ire, icp, -2, [30, 20, 1.0 / 920.0, 1.7, 1.2])
    assert cntc.gethertzcontact(ire, icp)[6] == 1.2
    cntc.sethertzcontact(ire, icp, -6, [11, 12, 5.0, 4.0, 1.5])       # wrong count: ignored
    assert cntc.gethertzcontact(ire, icp)[6] == 1.2
    cntc.sethertzcontact(ire, icp, -7, [13, 14, 5.0, 4.0, 1.6])       # no such method
    assert cntc.gethertzcontact(ire, icp)[6] == 1.2
    cntc.sethertzcontact(ire, icp, -6, [15, 16, 5.0, 4.0, 4.5, 1.7])  # accepted
    h = cntc.gethertzcontact(ire, icp)
    assert len(h) == 10 and h[6] == 1.7 and h[7] == 4.0 and h[8] == 4.5
    cntc.sethertzcontact(ire, icp, -5, [15, 16, 5.0, 4.0, 1.8])       # accepted, 5 parameters
    assert cntc.gethertzcontact(ire, icp)[6] == 1.8
    cntc.finalize(ire)


FLANG = next((p for p in ("/opt/rocm/bin/amdflang", "/opt/rocm/llvm/bin/flang") if os.path.exists(p)), None)


@pytest.mark.skipif(FLANG is None, reason="no Fortran compiler")
def test_fortran_caller_through_iso_c_binding(cntc, tmp_path):
    """The reference's host language drives the library: a Fortran program with ISO_C_BINDING
    interface blocks (tests/fortran/cntc_caller.f90) links libcntc_amd.so, issues the set-calls,
    reads them back and calls cntc_calculate, which without a GPU returns CNTC_err_other."""
    exe = tmp_path / "caller"
    src = os.path.join(ROOT, "tests", "fortran", "cntc_caller.f90")
    libdir = os.path.dirname(cntc.library_path)
    subprocess.run([FLANG, "-O1", "-o", str(exe), src, "-L" + libdir, "-lcntc_amd",
                    "-Wl,-rpath," + libdir], cwd=tmp_path, check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = dict(l.split(None, 1) for l in r.stdout.splitlines() if " " in l.strip())
    assert out["flags"].split() == ["3", "1", "0"]
    assert float(out["veloc"]) == 10000.0
    assert [float(x) for x in out["moduli"].split()] == [82000.0] * 3
    assert float(out["scale"]) == 1.3
    rc = int(out["ierror"])
    if os.path.exists("/dev/kfd"):
        assert rc == 0 and float(out["forces"].split()[0]) == pytest.approx(1e5, rel=1e-12)
    else:
        assert rc == -99
    assert "FORTRAN_CALLER_DONE" in r.stdout
