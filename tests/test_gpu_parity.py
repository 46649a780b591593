"""Parity of the HIP engine (through the cntc_* C ABI) with the CPU restatement.

Tolerances: element states identical; tractions and forces to 1e-8 relative of the force
scale when both solves run to eps=1e-10 or tighter; with the reference's own solver
settings (eps=1e-6 .. 1e-8), forces within 1e-6 relative and states identical except at
elements whose slip/traction margin lies within the solver tolerance.
"""
import math

import numpy as np
import pytest

from tests import cases as C

pytestmark = pytest.mark.gpu


def run_gpu(cntc, c, ire=1, icp=1):
    C.apply_cntc(cntc, ire, icp, c)
    rc = cntc.calculate(ire, icp)
    fn, fx, fy, mz = cntc.getcontactforces(ire, icp)
    pn, px, py = cntc.gettractions(ire, icp)
    el = cntc.getelementdivision(ire, icp)
    return rc, (fn, fx, fy, mz), (pn.ravel(), px.ravel(), py.ravel()), el.ravel()


def run_oracle(O, c):
    P = O.Problem()
    inp, h = C.to_oracle(O, c)
    rc = P.calculate(inp, h)
    return rc, P.forces(), P.tractions(), P.eldiv(), P


def compare(res_g, res_o, ftol, ptol, name, max_el_diff=0):
    rc_g, f_g, p_g, el_g = res_g
    rc_o, f_o, p_o, el_o, _ = res_o
    print(name, "gpu rc", rc_g, "forces", f_g, "| oracle rc", rc_o, "forces", f_o, flush=True)
    assert rc_g == rc_o == 0
    fscale = max(abs(f_o[0]), 1e-30)
    for a, b in zip(f_g[:3], f_o[:3]):
        assert abs(a - b) <= ftol * fscale, (name, f_g, f_o)
    pscale = max(np.abs(p_o[0]).max(), 1e-30)
    for a, b in zip(p_g, p_o):
        assert np.abs(a - b).max() <= ptol * pscale, (name, np.abs(a - b).max() / pscale)
    ndiff = int((el_g != el_o).sum())
    assert ndiff <= max_el_diff, (name, ndiff)


def test_cattaneo_tangcg(cntc, oracle):
    """testbank.inp case 1 (T=1, TangCG with FFT preconditioner)."""
    c = C.cattaneo_case()
    g = run_gpu(cntc, c)
    o = run_oracle(oracle, c)
    compare(g, o, 1e-8, 1e-7, "cattaneo")
    # golden, testbank/testbank.ref_out:44-47: FN 1720., FX -451.5, C=177, A=45, S=132
    fn, fx = g[1][0], g[1][1]
    assert abs(fn - 1720.) < 0.5 and abs(fx - (-451.5)) < 0.05
    assert (g[3] > 0).sum() == 177 and (g[3] == 1).sum() == 45 and (g[3] == 2).sum() == 132


def test_table_first_cases(cntc, oracle):
    """First rows of test_table (mx=11, steady rolling, GDsteady) against the oracle."""
    rows, _ = C.table_cases(11, 3220)
    P = oracle.Problem()
    for r in rows[:40:3] + rows[1150:1160:3]:
        base = C.hertz_case(r["aa"], r["bb"], 11, 11, 82000.0, fstat=0.3, eps=1e-6, maxgs=299, maxin=1)
        rc = P.calculate(C.to_oracle(oracle, base)[0])
        _, _, _, _, rho, cp = P.hertz()
        cksi = r["cmgn"] * 0.3 * math.cos(r["cang"]) * cp / rho
        ceta = r["cmgn"] * 0.3 * math.sin(r["cang"]) * cp / rho
        cphi = 0.3 * r["spin"] / rho
        c = C.hertz_case(r["aa"], r["bb"], 11, 11, 82000.0, cksi=cksi, ceta=ceta, cphi=cphi,
                         eps=1e-6, maxgs=299, maxin=1)
        g = run_gpu(cntc, c)
        o = run_oracle(oracle, c)
        compare(g, o, 1e-6, 1e-5, "table case %d" % r["icase"], max_el_diff=2)


def test_steady_rolling_40(cntc, oracle):
    """The 40x40 steady-rolling Hertz patch (Fn 100 kN, xi = -0.001, GDsteady) at eps 1e-12:
    Fn, Fx each to 1e-9 of itself (Fy, zero here, to 1e-9 of mu Fn), tractions to 1e-8 of
    each component's largest value, element states identical."""
    from tests.test_gpu_configs import traction_err
    c = C.hertz_case(7.193, 5.411, 40, 40, 1e5, cksi=-0.001, eps=1e-12, maxgs=2000, maxin=20,
                     scale=1.3)
    g = run_gpu(cntc, c)
    o = run_oracle(oracle, c)
    assert g[0] == o[0] == 0
    # Fn, Fx relative to themselves; Fy vanishes by symmetry (eta = phi = 0): relative to mu Fn
    den = np.array([abs(o[1][0]), abs(o[1][1]), 0.3 * o[1][0]])
    ef = np.abs(np.array(g[1][:3]) - np.array(o[1][:3])) / den
    ep = traction_err(g[2], o[2])
    print("hertz40 eps 1e-12: gpu", g[1], "oracle", o[1], "force err", ef, "traction err %.2e" % ep)
    assert ef.max() <= 1e-9 and ep <= 1e-8
    assert np.array_equal(g[3], o[3])


def test_batch_equals_single(cntc, oracle):
    """Cases solved in one batched launch give the same results as solved one by one."""
    cs = [C.hertz_case(4.0 + k, 6.0 + 0.5 * k, 12 + 3 * k, 10 + 2 * k, 5e4 + 1e4 * k,
                       cksi=-0.0005 * (k + 1), ceta=0.0003 * k, eps=1e-8, maxgs=999, maxin=20)
          for k in range(5)]
    single = []
    for k, c in enumerate(cs):
        single.append(run_gpu(cntc, c, ire=10 + k))
        assert single[-1][0] == 0
        compare(single[-1], run_oracle(oracle, c), 1e-6, 1e-5, "batch case %d" % k, max_el_diff=2)
    for k, c in enumerate(cs):
        C.apply_cntc(cntc, 20 + k, 1, c)
    err = cntc.calculatebatch([20 + k for k in range(5)], [1] * 5)
    assert (err == 0).all()
    for k in range(5):
        f = cntc.getcontactforces(20 + k, 1)
        pn, px, py = cntc.gettractions(20 + k, 1)
        assert f == single[k][1]
        assert np.array_equal(pn.ravel(), single[k][2][0])
        assert np.array_equal(px.ravel(), single[k][2][1])


def test_steadygs_table_golden(cntc, oracle):
    """test_table rows with the reference's default solver for T=3 (G=0: SteadyGS) against
    the golden rows of testbank/test_table_mx11.ref_fx and against the oracle."""
    rows, _ = C.table_cases(11, 3220)
    gold = C.table_golden()
    worst = 0.0
    for r in rows[::23]:
        c, cp = C.table_case_inputs(oracle, r, gausei=0)
        g = run_gpu(cntc, c)
        o = run_oracle(oracle, c)
        compare(g, o, 1e-6, 1e-5, "table G=0 case %d" % r["icase"], max_el_diff=2)
        fn, fx, fy, mz = g[1]
        el = g[3]
        vals = [fx / (0.3 * fn), fy / (0.3 * fn), mz / (0.3 * fn * cp)]
        ref = gold[r["icase"] - 1]
        d = max(abs(a - b) for a, b in zip(vals, ref[:3]))
        worst = max(worst, d)
        assert d <= 1.5e-6, (r["icase"], vals, ref)
        pct = 100.0 * (el == 2).sum() / (el > 0).sum()
        assert abs(pct - ref[3]) <= 0.051, (r["icase"], pct, ref[3])
    print("table G=0: worst golden difference %.2e" % worst)


@pytest.mark.parametrize("kernel", ["blocks", "overlapped", "staged", "direct"])
def test_gdsteady_stagnation_fallback(cntc, oracle, kernel, monkeypatch):
    """Hertzian patches on which GDsteady stagnates: both solvers fall back to SteadyGS
    (src/m_solvpt.f90:466-485) and must agree to the solver tolerance.  The four SteadyGS
    variants are checked: blocks of rows with the other rows' influence from FFT products (the
    default, rows up to ~650 elements), row sums overlapped with the walk along the previous
    row, row sums before each row's walk, and direct summation (rows too wide for staging)."""
    from contact_amd import workloads as W
    if kernel == "blocks":
        monkeypatch.setenv("CNTC_GS_KERNEL", "4")
    elif kernel == "direct":
        monkeypatch.setenv("CNTC_GS_KERNEL", "1")
    elif kernel == "staged":
        monkeypatch.setenv("CNTC_GS_KERNEL", "2")
    elif kernel == "overlapped":
        monkeypatch.setenv("CNTC_GS_KERNEL", "3")
    cases = W.hertz_batch(n=2048)
    for k in (381, 1505):
        c = cases[k]
        rc_g = run_gpu(cntc, c)
        its = cntc.getiterations(1, 1)
        o = run_oracle(oracle, c)
        print("patch", k, "gpu iterations", its, "oracle flags", o[4].stagnated())
        assert its[4] & 2, "SteadyGS did not run on the GPU"
        assert o[4].stagnated() & 2, "SteadyGS did not run in the oracle"
        compare(rc_g, o, 1e-6, 1e-5, "fallback patch %d" % k, max_el_diff=4)


def test_steadygs_wide_rows(cntc, oracle):
    """SteadyGS (G=0) on rows of 520 elements: the block variant then needs about 126 KB of LDS
    per work-group (gfx950 allows 160 KB)."""
    c = C.hertz_case(12.0, 3.0, 520, 16, 82000.0, cksi=0.0008, ceta=-0.0004, cphi=0.0002,
                     eps=1e-6, maxgs=299, maxin=5, gausei=0)
    g = run_gpu(cntc, c)
    its = cntc.getiterations(1, 1)
    o = run_oracle(oracle, c)
    print("wide rows: gpu iterations", its, "oracle", o[4].iters())
    assert its[4] & 2
    compare(g, o, 1e-6, 1e-5, "wide rows", max_el_diff=4)


def test_cattaneo_force_golden(cntc, oracle):
    """examples/cattaneo.inp case 2 through the C ABI (T=1, F=2: Fx, Fy prescribed, Newton on
    the shift): examples/cattaneo.ref_out:88-139 and the oracle."""
    c = C.cattaneo_force_case()
    g = run_gpu(cntc, c, ire=51)
    o = run_oracle(oracle, c)
    compare(g, o, 1e-8, 1e-7, "cattaneo F=2")
    cksi, ceta, _ = cntc.getcreepages(51, 1)
    ocksi, oceta, _ = o[4].creepages()
    print("creepages gpu", cksi, ceta, "oracle", ocksi, oceta, flush=True)
    assert abs(cksi - 8.155e-3) < 5e-7 and abs(ceta - (-6.716e-9)) < 5e-12
    assert abs(cksi - ocksi) <= 1e-10 * abs(ocksi) and abs(ceta - oceta) <= 1e-6 * abs(oceta)
    el = g[3]
    assert (el > 0).sum() == 177 and (el == 1).sum() == 45 and (el == 2).sum() == 132
    assert cntc.getiterations(51, 1)[3] == 59
    S = cntc.getsensitivities(51, 1)
    assert float("%.3g" % S[1, 1]) == -67.6 and float("%.3g" % S[2, 2]) == -117.0
    So = o[4].sens()
    assert np.allclose(S, So, rtol=1e-8, atol=0)


@pytest.mark.parametrize("force,gausei", [(1, 5), (2, 0), (2, 5)])
def test_force_prescribed_rolling(cntc, oracle, force, gausei):
    """T=3 with Fx (F=1) or Fx and Fy (F=2) prescribed: Newton on the creepages with GDsteady
    or SteadyGS inside, against the oracle; the creepages found and the sensitivities (S=3,
    whose right-hand side adds the solved creepages) agree too."""
    c = C.hertz_case(6.0, 4.0, 24, 24, 8200.0, cphi=0.0005, eps=1e-9, maxgs=2000, maxin=40,
                     gausei=gausei)
    c["ceta"] = 0.0002
    c.update(force=force, fxrel=-0.6, fyrel=0.3, sens=3)
    ire = 52 + force * 2 + (gausei == 0)
    g = run_gpu(cntc, c, ire=ire)
    o = run_oracle(oracle, c)
    compare(g, o, 1e-8, 1e-7, "rolling F=%d G=%d" % (force, gausei))
    cr = cntc.getcreepages(ire, 1)
    ocr = o[4].creepages()
    print("creepages gpu", cr, "oracle", ocr, flush=True)
    assert abs(cr[0] - ocr[0]) <= 1e-8 * abs(ocr[0]) and abs(cr[1] - ocr[1]) <= 1e-8 * max(abs(ocr[1]), 1e-6)
    S, So = cntc.getsensitivities(ire, 1), o[4].sens()
    for k in range(4):
        scale = np.abs(So[:, k]).max()
        assert np.abs(S[:, k] - So[:, k]).max() <= 1e-5 * scale, (k, S, So)
