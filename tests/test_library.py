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
