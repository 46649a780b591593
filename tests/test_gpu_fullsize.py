"""Full-size rough non-Hertzian patches (gap = Hertz quadratic + seeded Gaussian roughness,
RMS 1 um, correlation length 4 elements, Fn = 120 kN, xi = 0.002, eta = 0.001, phi = 0.1 rad/m):

* 256^2 and 512^2 in FP64 against the oracle's results stored in tests/golden (made by
  tests/golden/make_rough_fixture.py);
* 1024^2 with FP32 products in the FP64 refinement against the FP64 solve of the same patch.
"""
import os

import numpy as np
import pytest

from contact_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def solve(cntc, ire, c):
    W.apply_case(cntc, ire, 1, c)
    rc = cntc.calculate(ire, 1)
    f = np.array(cntc.getcontactforces(ire, 1))
    pn, px, py = (a.ravel() for a in cntc.gettractions(ire, 1))
    el = cntc.getelementdivision(ire, 1).ravel()
    return rc, f, pn, px, py, el


@pytest.mark.parametrize("n", [256, 512])
def test_rough_patch_fp64_vs_oracle(cntc, n):
    g = np.load(os.path.join(HERE, "golden", "rough_patch_%d.npz" % n))
    rc, f, pn, px, py, el = solve(cntc, 40 + n // 256, W.rough_patch(n=n))
    print("n", n, "gpu", f, cntc.getiterations(40 + n // 256, 1), "oracle", g["forces"], g["iters"])
    assert rc == 0
    fo = g["forces"]
    scale = np.linalg.norm(fo[:3])
    assert np.max(np.abs(f[:3] - fo[:3])) <= 1e-8 * scale
    s = g["stride"]
    pmax = g["sums"][3]
    for a, b in ((pn, g["pn"]), (px, g["px"]), (py, g["py"])):
        assert np.max(np.abs(a[::s] - b)) <= 1e-6 * pmax
    counts = np.bincount(el, minlength=4)
    assert np.abs(counts - g["counts"]).sum() <= 8, (counts, g["counts"])


def test_rough_patch_1024_mixed_vs_fp64(cntc):
    c = W.rough_patch(n=1024)
    r64 = solve(cntc, 50, c)
    cm = dict(c)
    cm["precision"] = 1
    rmx = solve(cntc, 51, cm)
    print("1024 fp64", r64[1], cntc.getiterations(50, 1), "mixed", rmx[1], cntc.getiterations(51, 1))
    assert r64[0] == 0 and rmx[0] == 0
    rel = np.abs(rmx[1][:3] - r64[1][:3]) / np.abs(r64[1][:3])
    pmax = np.max(r64[2])
    terr = max(np.max(np.abs(a - b)) for a, b in zip(rmx[2:5], r64[2:5])) / pmax
    print("force relative error", rel, "traction error / pmax", terr)
    assert np.max(rel) <= 1e-6
