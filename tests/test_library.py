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


_INP_SCRIPT = r'''
import ctypes, sys
sys.path.insert(0, sys.argv[1])
from contact_amd import cntc, workloads as W
i = ctypes.c_int
L = cntc.lib()
out, exp = sys.argv[2].encode(), b"replay/exp1"
L.cntc_initializefirst(ctypes.byref(i(0)), ctypes.byref(i(0)), ctypes.byref(i(0)), b" ", out, exp,
                       ctypes.byref(i(1)), ctypes.byref(i(len(out))), ctypes.byref(i(len(exp))))
c = W.base_case(ipotcn=-3, mx=40, my=40, aa=7.193, bb=5.411, scale=1.3, fn=1e5, cksi=-0.001,
                ceta=0.0, cphi=0.0, eps=1e-8, maxgs=999, maxin=20, fstat=0.3)
W.apply_case(cntc, 5, 1, c)
cntc.setflags(5, 1, [cntc.CNTC["if_wrtinp"]], [1])
print(cntc.calculate(5, 1))
W.apply_case(cntc, 6, 2, W.rough_patch(n=12))
cntc.setflags(6, 2, [cntc.CNTC["if_wrtinp"]], [1])
print(cntc.calculate(6, 2))
W.apply_case(cntc, 7, 1, c)      # no wrtinp flag: not written
print(cntc.calculate(7, 1))
'''


def test_inp_writer_replay_file(cntc, tmp_path):
    """CNTC_if_wrtinp: the case descriptions in <outdir>/<expnam>.inp in the layout of wrtinp
    (src/m_sinput.f90:1506-1862); written after the solve whatever its outcome (here without a
    GPU).  Checks the Fortran edit descriptors on known values (Gw.d switches between the F
    form with four trailing blanks and the E form at 0.1 and 10^d)."""
    from contact_amd import workloads as W
    r = subprocess.run([sys.executable, "-c", _INP_SCRIPT, ROOT, str(tmp_path)], capture_output=True,
                       text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    lines = (tmp_path / "exp1.inp").read_text().split("\n")
    # case 1: Hertzian 40 x 40 steady rolling with GDsteady (G=5), Fn prescribed
    assert lines[0] == "   3   5   1  MODULE, RE, CP: case 1 for result element  5, contact problem  1"
    assert lines[1] == "  203100          P-B-T-N-F-S           PVTIME, BOUND,  TANG,   NORM,   FORCE,  STRESS"
    assert lines[2] == "  022020          L-D-C-M-Z-E           FRCLAW, DISCNS, INFLCF, MATER,  RZNORM, EXRHS"
    assert lines[3] == " 0501241        H-G-I-A-O-W-R     HEAT, GAUSEI, IESTIM, MATFIL, OUTPUT, FLOW,   RETURN"
    assert lines[4] == "   999    20   30    1    1.0E-08        MAXGS , MAXIN , MAXNR , MAXOUT, EPS"
    g = W.GD_PARAMS
    f = lines[5].split()
    assert [float(x) for x in f[:2]] == [round(g["fdecay"], 3), round(g["betath"], 3)] and f[2] == str(g["kdowfb"])
    assert lines[6] == "  0.1000E+06 -0.1000E-02   0.000       0.000              FN, CKSI, CETA, CPHI"
    assert lines[7] == "  0.3000      0.3000    " + " " * 34 + "FSTAT, FKIN"
    chi, dq, veloc = (float(x) for x in lines[8].split()[:3])
    assert chi == 0.0 and dq > 0.0 and veloc > 0.0 and lines[8].endswith("CHI, DQ, VELOC")
    assert lines[9] == "  0.2800      0.2800      0.8200E+05  0.8200E+05          POISS 1,2,  GG 1,2"
    assert lines[10] == "   -3                                                     IPOTCN"
    assert lines[11] == "   40   40   7.193       5.411       1.300                MX,MY,AA,BB,SCALE"
    assert lines[12] == " "
    # case 2: an element-wise gap (ibase 9) on a 12 x 12 grid: no solution, so the tight box is
    # the centre of the grid (4..8), 5 values per row
    k = lines.index("   3   6   2  MODULE, RE, CP: case 1 for result element  6, contact problem  2")
    body = lines[k:]
    ip = body.index("    1                                                     IPOTCN")
    assert body[ip + 1].startswith("    5    5") and body[ip + 1].endswith("MX,MY,XL,YL,DX,DY")
    assert body[ip + 2] == "    9    1                                                IBASE, IPLAN"
    assert body[ip + 3] == "%  UNDEF.DISTANCE, (1)-(2): SPECIFIED PER ELEMENT"
    # rows 4..8, each after a blank record (the first row is not row 1)
    rows = body[ip + 4: ip + 14]
    assert all(x == " " for x in rows[0::2])
    assert all(len(x) == 70 and len(x.split()) == 5 for x in rows[1::2])
    assert body[ip + 14] == "%  UNRESTRICTED PLANFORM"
    # Gw.d on known values: E form below 0.1 (0.0338 -> 0.338080E-01), F form from 0.1
    assert rows[1].split()[0] == "0.338080E-01"
    # problem 7 did not ask for the file
    assert not any("result element  7" in x for x in lines)
