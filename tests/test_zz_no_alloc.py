"""The entry point that allocates no device memory (cntc_calculatenoalloc) against
cntc_calculate on the same cases: identical results, bit for bit, on two calls in a row (the
second with a smaller patch), on a SteadyGS case, over a transient-rolling continuation, and a
clear error for a patch larger than its fixed array."""
import numpy as np
import pytest

from tests import cases as C
from contact_amd import workloads as W

pytestmark = pytest.mark.gpu


def results(cntc, ire):
    f = cntc.getcontactforces(ire, 1)
    tr = [a.ravel().copy() for a in cntc.gettractions(ire, 1)]
    sl = [a.ravel().copy() for a in cntc.getmicroslip(ire, 1)]
    el = cntc.getelementdivision(ire, 1).ravel().copy()
    un = [a.ravel().copy() for a in cntc.getdisplacements(ire, 1)]
    return f, tr + sl + un, el


def same(a, b, name):
    assert a[0] == b[0], (name, a[0], b[0])
    for x, y in zip(a[1], b[1]):
        assert np.array_equal(x, y), name
    assert np.array_equal(a[2], b[2]), name


def both(cntc, c, ire_ref, ire_na, name):
    W.apply_case(cntc, ire_ref, 1, c)
    assert cntc.calculate(ire_ref, 1) == 0
    W.apply_case(cntc, ire_na, 1, c)
    rc = cntc.calculatenoalloc(ire_na, 1)
    assert rc == 0, (name, rc)
    same(results(cntc, ire_ref), results(cntc, ire_na), name)
    print(name, "forces", cntc.getcontactforces(ire_na, 1), flush=True)


def test_noalloc_equals_calculate(cntc, monkeypatch):
    for ire in range(90, 100):
        cntc.initialize(ire)
    # two calls in a row, the second with a smaller patch
    big = C.hertz_case(7.193, 5.411, 40, 40, 1e5, cksi=-0.001, ceta=0.0002, cphi=0.05e-3,
                       eps=1e-10, maxgs=2000, maxin=20, scale=1.3)
    both(cntc, big, 90, 91, "40x40 GDsteady")
    small = C.cattaneo_case()
    both(cntc, small, 92, 93, "17x17 TangCG")
    # SteadyGS (G=0), run on the main stream by this entry point with the one-work-group kernel
    # (row sums overlapped with the walk); cntc_calculate is pinned to the same kernel, since
    # its default (blocks of rows with FFT products) sums in another order
    gs = C.hertz_case(5.0, 3.5, 26, 20, 6e4, cksi=-0.0012, ceta=0.0003, eps=1e-8, maxgs=999,
                      maxin=20, gausei=0)
    monkeypatch.setenv("CNTC_GS_KERNEL", "3")
    both(cntc, gs, 94, 95, "26x20 SteadyGS")
    monkeypatch.delenv("CNTC_GS_KERNEL")
    # transient rolling: continuation and warm start from the solution kept on the host
    for step in range(3):
        c = C.hertz_case(6.0, 4.0, 30, 24, 8e4, cksi=-0.0005 * (step + 1), ceta=0.0001,
                         eps=1e-9, maxgs=500, maxin=50, tang=2)
        c["pvtime"] = 2 if step == 0 else 0
        c["iestim"] = 0 if step == 0 else 3
        both(cntc, c, 96, 97, "T=2 step %d" % step)


def test_noalloc_rejects_a_patch_too_large(cntc):
    cntc.initialize(98)
    c = C.hertz_case(9.0, 9.0, 200, 200, 1e5, cksi=-0.001, eps=1e-8, maxgs=999, maxin=20)
    W.apply_case(cntc, 98, 1, c)
    assert cntc.calculatenoalloc(98, 1) == cntc.CNTC["err_other"]
    # the problem itself is intact: the ordinary entry point solves it
    small = C.hertz_case(5.0, 4.0, 24, 20, 6e4, cksi=-0.001, eps=1e-8, maxgs=999, maxin=20)
    W.apply_case(cntc, 98, 1, small)
    assert cntc.calculatenoalloc(98, 1) == 0
    W.apply_case(cntc, 99, 1, small)
    assert cntc.calculate(99, 1) == 0
    assert cntc.getcontactforces(98, 1) == cntc.getcontactforces(99, 1)
