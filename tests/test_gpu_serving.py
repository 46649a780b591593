"""Transient rolling and the continuous-batching loop on the GPU.

* T=2 transient rolling (dq = dx, TangCG) over several steps of continuation with warm
  starts, against the oracle.
* Requests submitted while others are in flight give the same result, bit for bit, as the
  same case solved alone (per-patch arithmetic does not depend on the other patches in a
  launch).
"""
import time

import numpy as np
import pytest

from tests import cases as C
from contact_amd import workloads as W

pytestmark = pytest.mark.gpu


def transient_case(step, cksi):
    c = C.hertz_case(6.0, 4.0, 30, 24, 80000.0, cksi=cksi, ceta=-0.3 * cksi, cphi=0.0,
                     eps=1e-9, maxgs=500, maxin=50, tang=2)
    c["pvtime"] = 2 if step == 0 else 0
    c["iestim"] = 0 if step == 0 else 3
    return c


def test_transient_rolling_matches_oracle(cntc, oracle):
    P = oracle.Problem()
    for step, cksi in enumerate([0.0005, 0.0008, 0.0010, 0.0011, 0.0011]):
        c = transient_case(step, cksi)
        C.apply_cntc(cntc, 30, 1, c)
        rc = cntc.calculate(30, 1)
        fg = cntc.getcontactforces(30, 1)
        pn, px, py = (a.ravel() for a in cntc.gettractions(30, 1))
        el = cntc.getelementdivision(30, 1).ravel()
        inp, _ = C.to_oracle(oracle, c)
        rco = P.calculate(inp)
        fo = P.forces()
        opn, opx, opy = P.tractions()
        print("step", step, "gpu", rc, fg, "oracle", rco, fo, cntc.getiterations(30, 1), P.iters())
        assert rc == 0 and rco == 0
        for a, b in zip(fg[:3], fo[:3]):
            assert abs(a - b) <= 1e-8 * abs(fo[0])
        scale = np.abs(opn).max()
        for a, b in zip((pn, px, py), (opn, opx, opy)):
            assert np.abs(a - b).max() <= 1e-7 * scale
        assert int((el != P.eldiv()).sum()) <= 2


def _solve_alone(cntc, ire, icp, c):
    W.apply_case(cntc, ire, icp, c)
    rc = cntc.calculate(ire, icp)
    return rc, cntc.getcontactforces(ire, icp), [a.copy() for a in cntc.gettractions(ire, icp)]


def test_serving_results_equal_alone(cntc):
    cases = [c for c in W.hertz_batch(n=256) if c["mx"] * c["my"] <= 4000][:24]
    alone = [_solve_alone(cntc, 500 + k // 9, 1 + k % 9, c) for k, c in enumerate(cases)]
    ids = [(600 + k // 9, 1 + k % 9) for k in range(len(cases))]
    for (ire, icp), c in zip(ids, cases):
        W.apply_case(cntc, ire, icp, c)
    assert cntc.servestart(0.0) == 0
    tickets = []
    for k, (ire, icp) in enumerate(ids):
        t = cntc.submit(ire, icp)
        assert t > 0
        tickets.append(t)
        if k % 6 == 5:
            time.sleep(0.02)          # later requests arrive while earlier ones run
    lat = []
    for t in tickets:
        rc, l = cntc.wait(t)
        assert rc == 0
        lat.append(l)
    cntc.servestop()
    for k, (ire, icp) in enumerate(ids):
        f = cntc.getcontactforces(ire, icp)
        tr = cntc.gettractions(ire, icp)
        assert f == alone[k][1], (k, f, alone[k][1])
        for a, b in zip(tr, alone[k][2]):
            assert np.array_equal(a, b), k
    print("serving latencies (ms): p50 %.1f max %.1f" % (np.median(lat) * 1e3, max(lat) * 1e3))


def test_serving_channels_warm_start(cntc):
    """Channels stepping T=2 transient rolling through the loop, each step submitted when the
    channel's previous step has finished, against the same steps solved one by one."""
    base = [c for c in W.hertz_batch(n=64) if c["mx"] * c["my"] <= 3000][:4]
    steps = 3
    def case(c, s):
        d = dict(c)
        d.update(tang=2, eps=1e-8, maxgs=500, maxin=50, pvtime=2 if s == 0 else 0,
                 iestim=0 if s == 0 else 3, cksi=c["cksi"] * (1 + 0.1 * s))
        return d
    ref = []
    for ch, c in enumerate(base):
        out = []
        for s in range(steps):
            out.append(_solve_alone(cntc, 700, 1 + ch, case(c, s)))
            assert out[-1][0] == 0
        ref.append(out)
    cntc.servestart(0.0)
    nxt = [0] * len(base)
    pending = {}
    for ch, c in enumerate(base):
        W.apply_case(cntc, 710, 1 + ch, case(c, 0))
        pending[cntc.submit(710, 1 + ch)] = ch
    got = [[] for _ in base]
    while pending:
        for t in list(pending):
            done, rc, _ = cntc.poll(t)
            if done != 1:
                continue
            ch = pending.pop(t)
            assert rc == 0
            got[ch].append(cntc.getcontactforces(710, 1 + ch))
            nxt[ch] += 1
            if nxt[ch] < steps:
                W.apply_case(cntc, 710, 1 + ch, case(base[ch], nxt[ch]))
                pending[cntc.submit(710, 1 + ch)] = ch
        time.sleep(0.001)
    cntc.servestop()
    for ch in range(len(base)):
        for s in range(steps):
            assert got[ch][s] == ref[ch][s][1], (ch, s)


def test_transient_rolling_reaches_the_steady_state(cntc, oracle):
    """The GPU's T=2 path (continuation P=0, warm start I=3, rolled coefficients) stepped at fixed
    creepages reaches the steady-rolling (T=3, SteadyGS) solution of the oracle: the fixed point
    of transient rolling satisfies the steady-rolling equations."""
    base = C.hertz_case(6.0, 4.0, 30, 24, 80000.0, cksi=-0.0008, ceta=0.0003, eps=1e-12,
                        maxgs=2000, maxin=50, gausei=0)
    P3 = oracle.Problem()
    assert P3.calculate(*C.to_oracle(oracle, base)) == 0
    f3, p3, el3 = np.array(P3.forces()), P3.tractions(), P3.eldiv()
    dist = []
    for step in range(450):
        c = dict(base, tang=2, pvtime=2 if step == 0 else 0, iestim=0 if step == 0 else 3)
        C.apply_cntc(cntc, 31, 1, c)
        assert cntc.calculate(31, 1) == 0
        f = np.array(cntc.getcontactforces(31, 1))
        dist.append(np.abs(f[:3] - f3[:3]).max() / abs(f3[1]))
    print("GPU T=2 distance to the oracle's T=3 solution after 1, 100, 200, 450 steps:",
          dist[0], dist[99], dist[199], dist[-1])
    assert dist[0] > 1e-2 and dist[99] > 1e-5
    assert dist[-1] <= 1e-10
    p = [a.ravel() for a in cntc.gettractions(31, 1)]
    assert max(np.abs(a - b).max() for a, b in zip(p, p3)) <= 1e-9 * np.abs(p3[0]).max()
    assert np.array_equal(cntc.getelementdivision(31, 1).ravel(), el3)
