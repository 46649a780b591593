"""Full-size rough non-Hertzian patches (gap = Hertz quadratic + seeded Gaussian roughness,
RMS 1 um, correlation length 4 elements, Fn = 120 kN, xi = 0.002, eta = 0.001, phi = 0.1 rad/m):

* 256^2 in FP64 at eps 1e-12 against the oracle's full solution at eps 1e-12, and 512^2 at
  eps 1e-12 against the oracle's forces and a traction sample (tests/golden, made by
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


def test_rough_patch_256_tight_vs_oracle(cntc):
    """256^2 at eps 1e-12 against the oracle at eps 1e-12, every element: forces to 1e-9 of each
    component, tractions to 1e-8 of each component's largest value, element states identical
    except where an element lies on a state boundary within 1e-8 of the largest pressure."""
    from tests.test_gpu_configs import force_err, traction_err, states_agree
    g = np.load(os.path.join(HERE, "golden", "rough_patch_256_eps12.npz"))
    c = W.rough_patch(n=256, eps=1e-12)
    rc, f, pn, px, py, el = solve(cntc, 601, c)
    assert rc == 0 and int(g["rc"]) == 0
    ref = [g["pn"], g["px"], g["py"]]
    ef = force_err(f, g["forces"], c["fstat"], c["fn"])
    ep = traction_err([pn, px, py], ref)
    ndiff, bad = states_agree(el, g["el"], [pn, px, py], ref, c["fstat"], 1e-8)
    print("256^2 eps 1e-12: gpu", f, cntc.getiterations(601, 1), "oracle", g["forces"], g["iters"],
          "force err", ef, "traction err %.2e" % ep, "states differing", ndiff)
    assert ef.max() <= 1e-9
    assert ep <= 1e-8
    assert not bad, bad[:10]


def test_rough_patch_512_tight_vs_oracle(cntc):
    """512^2 at eps 1e-12 against the oracle at eps 1e-12: forces to 1e-9 of each component,
    tractions to 1e-8 on a sample of every 97th element, element counts and the sampled
    states identical except on a state boundary within 1e-8 of the largest pressure."""
    from tests.test_gpu_configs import force_err, traction_err, states_agree
    g = np.load(os.path.join(HERE, "golden", "rough_patch_512_eps12.npz"))
    c = W.rough_patch(n=512, eps=1e-12)
    rc, f, pn, px, py, el = solve(cntc, 602, c)
    assert rc == 0 and int(g["rc"]) == 0
    s = int(g["stride"])
    got = [pn[::s], px[::s], py[::s]]
    ref = [g["pn"], g["px"], g["py"]]
    ef = force_err(f, g["forces"], c["fstat"], c["fn"])
    ep = traction_err(got, ref)
    ndiff, bad = states_agree(el[::s], g["el"], got, ref, c["fstat"], 1e-8)
    counts = np.bincount(el, minlength=4)
    print("512^2 eps 1e-12: gpu", f, cntc.getiterations(602, 1), "oracle", g["forces"], g["iters"],
          "force err", ef, "traction err %.2e" % ep, "counts", counts, g["counts"])
    assert ef.max() <= 1e-9
    assert ep <= 1e-8
    assert not bad, bad[:10]
    assert np.abs(counts - g["counts"]).sum() <= 4, (counts, g["counts"])


def test_rough_patch_1024_mixed_vs_fp64(cntc):
    c = W.rough_patch(n=1024)
    r64 = solve(cntc, 603, c)
    cm = dict(c)
    cm["precision"] = 1
    rmx = solve(cntc, 604, cm)
    print("1024 fp64", r64[1], cntc.getiterations(603, 1), "mixed", rmx[1], cntc.getiterations(604, 1))
    assert r64[0] == 0 and rmx[0] == 0
    rel = np.abs(rmx[1][:3] - r64[1][:3]) / np.abs(r64[1][:3])
    pmax = np.max(r64[2])
    terr = max(np.max(np.abs(a - b)) for a, b in zip(rmx[2:5], r64[2:5])) / pmax
    print("force relative error", rel, "traction error / pmax", terr)
    assert np.max(rel) <= 1e-6
