"""The benchmark's configurations run through the HIP library on their full workloads.

* 2048-patch Hertzian batch (sides log-uniform 20-256, seed 2024), FP64, all patches submitted
  at once to the continuous-batching loop: every patch solves; a size-stratified subset gives
  the same bits when solved alone; the patches of tests/golden/batch_subset.npz match the
  oracle (forces within the solver tolerance at eps 1e-8; forces to 1e-9 and tractions to
  1e-8 at eps 1e-12, element states identical except within the solver margin).
* The same batch with FP32 FFT products inside the FP64 refinement: every patch's forces
  within 1e-6 of the FP64 batch, and the subset within 1e-6 of the oracle.
* Serving: 64 wheel channels stepping T=2 transient rolling (warm-started from their previous
  tractions), Poisson arrivals offered at 1000 requests/s, 10% asking sensitivities; every
  result (forces, tractions, sensitivities) equals the same sequence solved alone.

Tolerance conventions (written into the helpers below): a force component F is compared
relative to max(|F_ref|, 1e-4 mu Fn) -- a component that vanishes (no lateral creepage)
has no relative error of its own; a traction component relative to the largest |value| of
that component in the reference, floored at 1e-3 of the largest pressure.
"""
import os
import time

import numpy as np
import pytest

from contact_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
NPATCH = 2048


def ids_of(base, k):
    return base + k // 9, 1 + k % 9


def force_err(f, ref, mu, fn):
    """Per-component (Fn, Fx, Fy) error relative to max(|ref|, 1e-4 mu Fn)."""
    f, ref = np.asarray(f, float)[:3], np.asarray(ref, float)[:3]
    den = np.maximum(np.abs(ref), 1e-4 * mu * abs(fn))
    return np.abs(f - ref) / den


def traction_err(p, ref):
    """max over (pn, px, py) of |p - ref| / max(max|ref_c|, 1e-3 max pn)."""
    pmax = np.abs(ref[0]).max()
    return max(np.abs(a - b).max() / max(np.abs(b).max(), 1e-3 * pmax) for a, b in zip(p, ref))


def states_agree(el, el_ref, p, ref, mu, tol):
    """Element states identical except where the element lies on a state boundary within
    `tol` of the largest pressure: contact/exterior with pn ~ 0, adhesion/slip with |pt| ~ mu pn."""
    d = np.nonzero(el != el_ref)[0]
    pmax = np.abs(ref[0]).max()
    bad = []
    for i in d:
        pn = max(abs(p[0][i]), abs(ref[0][i]))
        pt = np.hypot(ref[1][i], ref[2][i])
        on_ce = pn <= tol * pmax
        on_as = abs(pt - mu * ref[0][i]) <= tol * mu * pmax
        if not (on_ce or on_as):
            bad.append(int(i))
    return len(d), bad


def solve_alone(cntc, ire, icp, c):
    W.apply_case(cntc, ire, icp, c)
    rc = cntc.calculate(ire, icp)
    return rc, cntc.getcontactforces(ire, icp), [a.ravel().copy() for a in cntc.gettractions(ire, icp)]


def run_batch(cntc, cases, base):
    """All cases submitted at once to the serving loop; (error codes, forces, iteration rows)."""
    idl = [ids_of(base, k) for k in range(len(cases))]
    for (i, j), c in zip(idl, cases):
        W.apply_case(cntc, i, j, c)
    assert cntc.servestart(0.0) == 0
    t0 = time.perf_counter()
    tickets = [cntc.submit(i, j) for i, j in idl]
    assert min(tickets) > 0
    rc = np.array([cntc.wait(t)[0] for t in tickets])
    wall = time.perf_counter() - t0
    cntc.servestop()
    forces = np.array([cntc.getcontactforces(i, j) for i, j in idl])
    its = np.array([cntc.getiterations(i, j) for i, j in idl])
    print("batch of %d (ire %d..): %.1f s, %d SteadyGS fall-backs, %d with priority flushes"
          % (len(cases), base, wall, np.sum((its[:, 4] & 2) != 0), np.sum((its[:, 4] & 4) != 0)),
          flush=True)
    return rc, forces, its, idl


@pytest.fixture(scope="module")
def batch_fp64(cntc):
    cases = W.hertz_batch(n=NPATCH, seed=2024, eps=1e-8, precision=0)
    rc, forces, its, idl = run_batch(cntc, cases, 100)
    yield cases, rc, forces, its, idl
    for ire in range(100, 100 + NPATCH // 9 + 1):
        cntc.finalize(ire)


@pytest.fixture(scope="module")
def fixture_subset():
    return np.load(os.path.join(HERE, "golden", "batch_subset.npz"))


def test_batch_fp64_every_patch_solves(batch_fp64):
    cases, rc, forces, its, _ = batch_fp64
    assert np.all(rc == 0), sorted(set(rc.tolist()))
    assert np.all(np.isfinite(forces))
    fn = np.array([c["fn"] for c in cases])
    assert np.allclose(forces[:, 0], fn, rtol=1e-12)      # N=1: Fn as prescribed
    # the largest patches of the batch ran with flushes of their own, and the patches on which
    # GDsteady stagnates fell back to SteadyGS
    assert np.sum((its[:, 4] & 4) != 0) >= 1
    assert its[381, 4] & 2 and its[1505, 4] & 2


def test_batch_fp64_equals_alone(cntc, batch_fp64):
    """A size-stratified subset (ranks spread over the element counts, with the two
    stagnating patches and the largest patch with priority flushes) solved alone gives the
    batch's bits: forces and tractions."""
    cases, rc, forces, its, idl = batch_fp64
    n = np.array([c["mx"] * c["my"] for c in cases])
    order = np.argsort(n, kind="stable")
    pick = {int(order[i]) for i in np.linspace(0, NPATCH - 1, 24).round().astype(int)}
    prio = [k for k in np.argsort(-n) if its[k, 4] & 4]
    pick |= {381, 1505} | set(int(k) for k in prio[:2])
    for m, k in enumerate(sorted(pick)):
        r = solve_alone(cntc, *ids_of(560, m), cases[k])
        assert r[0] == 0
        assert r[1] == tuple(forces[k]), (k, r[1], forces[k])
        tb = [a.ravel() for a in cntc.gettractions(*idl[k])]
        for a, b in zip(r[2], tb):
            assert np.array_equal(a, b), k


def test_batch_fp64_subset_vs_oracle(cntc, batch_fp64, fixture_subset):
    """The fixture patches: the batch's forces against the oracle at the same eps (1e-8),
    within the solver tolerance; then the same patches solved at eps 1e-12 on the GPU against
    the oracle at eps 1e-12: forces to 1e-9, tractions to 1e-8, states identical except within
    the solver margin."""
    cases, rc, forces, its, idl = batch_fp64
    g = fixture_subset
    worst8 = worst12f = worst12p = 0.0
    for m, k in enumerate(g["patches"]):
        k = int(k)
        c = cases[k]
        assert int(g["rc_e8_%d" % k]) == 0
        e8 = force_err(forces[k], g["forces_e8_%d" % k], c["fstat"], c["fn"]).max()
        worst8 = max(worst8, e8)
        assert e8 <= 1e-6, (k, forces[k], g["forces_e8_%d" % k])
        c12 = dict(c)
        c12["eps"] = 1e-12
        r = solve_alone(cntc, *ids_of(630, m), c12)
        assert r[0] == 0
        ref = [g["pn_%d" % k], g["px_%d" % k], g["py_%d" % k]]
        ef = force_err(r[1], g["forces_e12_%d" % k], c["fstat"], c["fn"]).max()
        ep = traction_err(r[2], ref)
        el = cntc.getelementdivision(*ids_of(630, m)).ravel()
        ndiff, bad = states_agree(el, g["el_%d" % k], r[2], ref, c["fstat"], 1e-8)
        print("patch %d (%dx%d): e8 force %.1e; e12 force %.1e traction %.1e, %d states differ"
              % (k, c["mx"], c["my"], e8, ef, ep, ndiff), flush=True)
        worst12f, worst12p = max(worst12f, ef), max(worst12p, ep)
        assert ef <= 1e-9, (k, r[1], g["forces_e12_%d" % k])
        assert ep <= 1e-8, k
        assert not bad, (k, bad[:10])
    print("subset: worst force error %.2e (eps 1e-8); %.2e forces, %.2e tractions (eps 1e-12)"
          % (worst8, worst12f, worst12p))


def test_batch_mixed_precision_vs_fp64(cntc, batch_fp64, fixture_subset):
    """The whole batch with FP32 products inside the FP64 refinement: every patch's forces
    within 1e-6 of the FP64 batch, and the fixture patches within 1e-6 of the oracle."""
    cases64, rc64, f64, _, _ = batch_fp64
    cases = W.hertz_batch(n=NPATCH, seed=2024, eps=1e-8, precision=1)
    rc, fm, its, _ = run_batch(cntc, cases, 330)
    try:
        assert np.all(rc == 0), sorted(set(rc.tolist()))
        err = np.array([force_err(fm[k], f64[k], cases[k]["fstat"], cases[k]["fn"]).max()
                        for k in range(NPATCH)])
        raw = np.abs(fm[:, :3] - f64[:, :3]) / np.maximum(np.abs(f64[:, :3]), 1e-300)
        print("mixed vs FP64: max %.2e, p99 %.2e (unfloored relative: max %.2e)"
              % (err.max(), np.percentile(err, 99), raw.max()))
        assert err.max() <= 1e-6, (int(np.argmax(err)), err.max())
        g = fixture_subset
        for k in g["patches"]:
            k = int(k)
            e = force_err(fm[k], g["forces_e8_%d" % k], cases[k]["fstat"], cases[k]["fn"]).max()
            assert e <= 1e-6, (k, e)
    finally:
        for ire in range(330, 330 + NPATCH // 9 + 1):
            cntc.finalize(ire)


def test_serving_channels_at_1000_per_second(cntc):
    """64 channels (grids from the batch's size distribution) stepping T=2 transient rolling
    (dq = dx, continuation P=0 and warm start I=3 after the first step), requests as Poisson
    arrivals offered at 1000/s for 0.5 s, creepages as seeded random walks, 10% asking the
    sensitivities (S=3).  A channel's request waits for its previous one.  Every result --
    forces, tractions and, where asked, the 4x4 sensitivities -- equals the same channel's
    sequence solved alone, step by step."""
    nchan, rate, seconds = 64, 1000.0, 0.5
    rng = np.random.default_rng(77)
    chans = W.hertz_batch(n=nchan, seed=77)
    served = [ids_of(610, k) for k in range(nchan)]
    alone = [ids_of(620, k) for k in range(nchan)]
    state = []
    for (i, j), c in zip(served, chans):
        c = dict(c)
        c.update(tang=2, eps=1e-8, maxgs=500, maxin=50, pvtime=2, iestim=0)
        W.apply_case(cntc, i, j, c)
        state.append(c)
    err = cntc.calculatebatch([i for i, _ in served], [j for _, j in served])
    assert np.all(err == 0)
    for i, j in served:
        cntc.setflags(i, j, [1970, 1983], [0, 3])        # continuation, warm start
    n_req = int(rate * seconds)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    chan_of = rng.integers(0, nchan, n_req)
    walk = rng.normal(0.0, 2e-4, (n_req, 2))
    want_sens = rng.random(n_req) < 0.1
    log = [[] for _ in range(nchan)]        # per channel: (cksi, ceta, sens, forces, tractions, S)
    queue = [[] for _ in range(nchan)]
    inflight = {}
    assert cntc.servestart(0.0) == 0

    def start(ch, r):
        c = state[ch]
        c["cksi"] = float(np.clip(c["cksi"] + walk[r, 0], -0.005, 0.005))
        c["ceta"] = float(np.clip(c["ceta"] + walk[r, 1], -0.005, 0.005))
        i, j = served[ch]
        cntc.setcreepages(i, j, c["cksi"], c["ceta"], c["cphi"])
        cntc.setflags(i, j, [1988], [3 if want_sens[r] else 1])
        t = cntc.submit(i, j)
        assert t > 0
        inflight[t] = (ch, r, c["cksi"], c["ceta"])

    t0, nxt, done, lat = time.perf_counter(), 0, 0, []
    while done < n_req:
        now = time.perf_counter() - t0
        while nxt < n_req and arrivals[nxt] <= now:
            ch = int(chan_of[nxt])
            if any(v[0] == ch for v in inflight.values()):
                queue[ch].append(nxt)
            else:
                start(ch, nxt)
            nxt += 1
        for t in list(inflight):
            d, rc, l = cntc.poll(t)
            if d != 1:
                continue
            ch, r, ck, ce = inflight.pop(t)
            assert rc == 0
            lat.append(l)
            i, j = served[ch]
            S = cntc.getsensitivities(i, j) if want_sens[r] else None
            log[ch].append((ck, ce, bool(want_sens[r]), cntc.getcontactforces(i, j),
                            [a.ravel().copy() for a in cntc.gettractions(i, j)], S))
            done += 1
            if queue[ch]:
                start(ch, queue[ch].pop(0))
        time.sleep(0.0002)
    wall = time.perf_counter() - t0
    cntc.servestop()
    print("serving: %d requests offered at %.0f/s, achieved %.0f/s, latency p50 %.1f ms p99 %.1f ms"
          % (n_req, rate, n_req / wall, np.percentile(lat, 50) * 1e3, np.percentile(lat, 99) * 1e3))
    # the same sequences alone, step by step
    nsens = 0
    for ch in range(nchan):
        i, j = alone[ch]
        c = dict(chans[ch])
        c.update(tang=2, eps=1e-8, maxgs=500, maxin=50, pvtime=2, iestim=0)
        W.apply_case(cntc, i, j, c)
        assert cntc.calculate(i, j) == 0
        cntc.setflags(i, j, [1970, 1983], [0, 3])
        for ck, ce, sens, f, tr, S in log[ch]:
            cntc.setcreepages(i, j, ck, ce, c["cphi"])
            cntc.setflags(i, j, [1988], [3 if sens else 1])
            assert cntc.calculate(i, j) == 0
            assert cntc.getcontactforces(i, j) == f, ch
            for a, b in zip(cntc.gettractions(i, j), tr):
                assert np.array_equal(a.ravel(), b), ch
            if sens:
                nsens += 1
                assert np.array_equal(cntc.getsensitivities(i, j), S), ch
    assert nsens >= 10
    for ire in range(610, 630):
        cntc.finalize(ire)
