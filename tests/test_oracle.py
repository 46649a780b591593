"""CPU checks of the oracle (the C restatement in oracle/) against the reference's golden
outputs, closed forms and brute force.  No GPU is used.

Golden data: testbank/test_table_mx11.ref_fx (tests/golden/test_table_mx11_fx.csv),
testbank/test_table_mx11.ref_out:1-5 (Hertz rho, cp), testbank/testbank.ref_out:44-47
(Cattaneo case 1), examples/cattaneo.ref_out:88-139 (Cattaneo with prescribed forces: the
Newton iteration on the creepages, F=2).
"""
import math

import numpy as np
import pytest
from scipy import special

from tests import cases as C


# ---- opt_fft_size (src/m_aijpj.f90:1022-1119) ----------------------------------------------

def test_opt_fft_size_known_answers(oracle):
    # Hand-traced through the reference's pattern trading (start at the next power of two,
    # apply the patterns in order while the product stays >= n):
    #   40:  64 -(4->3)-> 48 -(6->5)-> 40
    #   80: 128 -(4->3)-> 96 -(32->27)-> 81   (not the smallest 7-smooth number, 80)
    #   71: 128 -(4->3)-> 96 -(4->3)-> 72
    #  512: 512 (every pattern would go below 512)
    known = {40: 40, 80: 81, 71: 72, 512: 512, 1: 1, 2: 2, 3: 3, 163: 168}
    for n, m in known.items():
        assert oracle.opt_fft_size(n) == m, (n, oracle.opt_fft_size(n), m)


def test_opt_fft_size_properties(oracle):
    def smooth7(v):
        for p in (2, 3, 5, 7):
            while v % p == 0:
                v //= p
        return v == 1
    for n in range(1, 1300):
        m = oracle.opt_fft_size(n)
        assert m >= n and smooth7(m) and m <= 2 ** math.ceil(math.log2(n)), (n, m)


# ---- FFT and the convolution (src/m_aijpj.f90:712-1015) -----------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 12, 13, 17, 30, 49, 97, 128, 210])
def test_fft_matches_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(oracle.fft(x, -1), np.fft.fft(x), rtol=0, atol=1e-12 * n)
    assert np.allclose(oracle.fft(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-12 * n)


@pytest.mark.parametrize("mx,my", [(5, 7), (11, 11), (16, 9), (13, 4), (1, 6)])
def test_conv_fft_equals_direct_sum(oracle, mx, my):
    rng = np.random.default_rng(mx * 100 + my)
    cf = rng.standard_normal((2 * my, 2 * mx))
    p = rng.standard_normal(mx * my)
    u_fft = oracle.conv_block(cf, p, method=1)
    u_dir = oracle.conv_block(cf, p, method=0)
    # brute force from the definition u(i) = sum_j cf(i - j) p(j)
    u_bf = np.zeros(mx * my)
    for iy in range(my):
        for ix in range(mx):
            s = 0.0
            for jy in range(my):
                for jx in range(mx):
                    s += cf[iy - jy + my, ix - jx + mx] * p[jy * mx + jx]
            u_bf[iy * mx + ix] = s
    assert np.allclose(u_dir, u_bf, rtol=0, atol=1e-12)
    assert np.allclose(u_fft, u_bf, rtol=0, atol=1e-11 * np.abs(u_bf).max())


def test_preconditioner_inverts_circulant(oracle):
    """fft_makePrec (src/m_aijpj.f90:457-708): M = G^2 IFFT(1/FFT(c)) over the periodic
    2mx x 2my grid, so the periodic convolution c * M equals G^2 times a unit impulse."""
    mx, my, ga = 6, 5, 82000.0
    cf = oracle.elascf_pcwcns(0.0, 0.28, mx, my, 0.5, 0.7)
    c = cf[0, 0]
    m = oracle.make_prec_block(c, ga)
    # both blocks are stored with offset (0, 0) at [my, mx]: shift to the periodic origin
    cs = np.roll(np.roll(c, -my, axis=0), -mx, axis=1)
    ms = np.roll(np.roll(m, -my, axis=0), -mx, axis=1)
    prod = np.real(np.fft.ifft2(np.fft.fft2(cs) * np.fft.fft2(ms)))
    delta = np.zeros_like(prod)
    delta[0, 0] = ga * ga
    assert np.allclose(prod, delta, rtol=0, atol=1e-9 * ga * ga)


# ---- influence coefficients (src/m_visc.f90:431-604) --------------------------------------

def test_coefficients_closed_forms(oracle):
    nu, mx, my = 0.28, 41, 41
    cf = oracle.elascf_pcwcns(0.0, nu, mx, my, 1.0, 1.0)

    def at(ik, jk, ix, iy):
        return cf[ik, jk, iy + my, ix + mx]

    # centre of a uniformly loaded unit square, two identical half-spaces (Boussinesq):
    # w = (1 - nu)/(pi G) * 4 d ln(1 + sqrt 2)
    assert at(2, 2, 0, 0) == pytest.approx((1 - nu) / math.pi * 4 * math.log(1 + math.sqrt(2)),
                                           rel=1e-13)
    # far field of a point load (Boussinesq, Cerruti), relative difference O((d/r)^2)
    for (ix, iy) in [(40, 0), (0, 40), (30, 30), (-25, 33)]:
        r = math.hypot(ix, iy)
        assert at(2, 2, ix, iy) == pytest.approx((1 - nu) / (math.pi * r), rel=1e-3)
        assert at(0, 0, ix, iy) == pytest.approx(((1 - nu) + nu * ix * ix / r ** 2) / (math.pi * r), rel=1e-3)
        assert at(1, 1, ix, iy) == pytest.approx(((1 - nu) + nu * iy * iy / r ** 2) / (math.pi * r), rel=1e-3)
        assert at(0, 1, ix, iy) == pytest.approx(nu * ix * iy / (math.pi * r ** 3), rel=1e-3, abs=1e-9)
    # symmetries: cf12 = cf21; even in x and y for the diagonal blocks; K = 0 decouples
    assert np.array_equal(cf[0, 1], cf[1, 0])
    for ik in range(3):
        assert np.allclose(at(ik, ik, 7, 3), at(ik, ik, -7, 3)) and np.allclose(at(ik, ik, 7, 3), at(ik, ik, 7, -3))
    assert np.all(cf[0, 2] == 0) and np.all(cf[2, 0] == 0) and np.all(cf[1, 2] == 0)


# ---- Hertz (src/m_hertz.f90:267-499) -------------------------------------------------------

def test_elliptic_integrals_vs_scipy(oracle):
    for mc in [1.0, 0.9, 0.5, 0.2, 0.05, 1e-3]:
        m = 1.0 - mc
        b, d = oracle.ellip_bd(mc)
        if m == 0.0:
            assert b == pytest.approx(math.pi / 4, rel=1e-14) and d == pytest.approx(math.pi / 4, rel=1e-14)
            continue
        K, E = special.ellipk(m), special.ellipe(m)
        assert b == pytest.approx((E - mc * K) / m, rel=1e-12)
        assert d == pytest.approx((K - E) / m, rel=1e-12)


def test_hertz_table_ellipses(oracle):
    """rho and cp of the five test_table ellipses (testbank/test_table_mx11.ref_out:1-5) and
    the curvatures from an independent solution with scipy's elliptic integrals."""
    from contact_amd.workloads import hertz_semiaxes
    printed = [(1772., 8.944), (2075., 8.485), (2700., 9.000), (2075., 8.485), (1772., 8.944)]
    ellip = [0.2, 0.5, 1.0, 2.0, 5.0]
    aa = [4.0, 6.0, 9.0, 12.0, 20.0]
    P = oracle.Problem()
    for k in range(5):
        c = C.hertz_case(aa[k], aa[k] / ellip[k], 11, 11, 82000.0)
        assert P.calculate(C.to_oracle(oracle, c)[0]) == 0
        a1, b1, ha, hb, rho, cp = P.hertz()
        assert round(rho) == printed[k][0] and round(cp, 3) == pytest.approx(printed[k][1], abs=1e-3)
        # forward check: the semi-axes of the gap a1 x^2 + b1 y^2 at Fn = 82 kN (the inverse
        # problem is solved iteratively, to about 1e-8)
        sa, sb = hertz_semiaxes(a1, b1, 82000.0, g=82000.0, nu=0.28)
        assert sa == pytest.approx(aa[k], rel=1e-7) and sb == pytest.approx(aa[k] / ellip[k], rel=1e-7)
    # circle: a = (3 P R / 4 E*)^(1/3), a1 = 1/(2R)
    c = C.hertz_case(9.0, 9.0, 11, 11, 82000.0)
    P.calculate(C.to_oracle(oracle, c)[0])
    a1 = P.hertz()[0]
    estar = 82000.0 / (1 - 0.28)
    assert a1 == pytest.approx(3 * 82000.0 / (8 * estar * 9.0 ** 3), rel=1e-12)


# ---- complete solves against the reference's golden outputs ----------------------------------

def _table_rows(oracle, gausei, step, tol=1.0e-6):
    rows, _ = C.table_cases(11, 3220)
    gold = C.table_golden()
    P = oracle.Problem()
    worst, nbad = 0.0, []
    for r in rows[::step]:
        c, cp = C.table_case_inputs(oracle, r, gausei)
        inp, _ = C.to_oracle(oracle, c)
        assert P.calculate(inp) == 0
        fn, fx, fy, mz = P.forces()
        el = P.eldiv()
        vals = np.array([fx / (0.3 * fn), fy / (0.3 * fn), mz / (0.3 * fn * cp)])
        ref = gold[r["icase"] - 1]
        d = float(np.max(np.abs(vals - ref[:3])))
        pct = 100.0 * (el == 2).sum() / (el > 0).sum()
        worst = max(worst, d)
        if d > tol + 1e-12 or abs(pct - ref[3]) > 0.051:
            nbad.append((r["icase"], vals, ref))
    return worst, nbad


def test_table_golden_steadygs(oracle):
    """All 3220 rows of test_table with its own solver (G=0 -> SteadyGS), printed to 6 digits."""
    worst, bad = _table_rows(oracle, 0, 1)
    assert not bad, bad[:5]
    assert worst <= 1.0e-6 + 1e-12


def test_table_golden_gdsteady(oracle):
    """Every 7th row with GDsteady (G=5, with the SteadyGS fall-back) instead.  The golden
    rows come from SteadyGS; both stop at eps = 1e-6, so they agree to 2e-6."""
    worst, bad = _table_rows(oracle, 5, 7, tol=2.0e-6)
    assert not bad, bad[:5]


def test_table_golden_detects_planted_error(oracle):
    """The golden comparison has teeth: a sign error in the spin creepage fails it."""
    rows, _ = C.table_cases(11, 3220)
    gold = C.table_golden()
    r = rows[23 * 7 * 1 + 5]              # a row with spin and creepage
    assert r["spin"] != 0.0
    c, cp = C.table_case_inputs(oracle, r, 0)
    c["cphi"] = -c["cphi"]
    P = oracle.Problem()
    P.calculate(C.to_oracle(oracle, c)[0])
    fn, fx, fy, mz = P.forces()
    vals = np.array([fx / (0.3 * fn), fy / (0.3 * fn), mz / (0.3 * fn * cp)])
    assert np.max(np.abs(vals - gold[r["icase"] - 1][:3])) > 1e-3


def test_cattaneo_golden(oracle):
    """testbank.inp case 1 (T=1 shift, TangCG): testbank/testbank.ref_out:44-47."""
    P = oracle.Problem()
    inp, h = C.to_oracle(oracle, C.cattaneo_case())
    assert P.calculate(inp, h) == 0
    fn, fx, fy, mz = P.forces()
    el = P.eldiv()
    assert abs(fn - 1720.) < 0.5 and abs(fx - (-451.5)) < 0.05
    assert (el > 0).sum() == 177 and (el == 1).sum() == 45 and (el == 2).sum() == 132


def _sig(x, ref, digits=4):
    """x printed with `digits` significant digits equals ref."""
    if ref == 0.0:
        return abs(x) < 5e-7
    return float("%.*e" % (digits - 1, x)) == pytest.approx(ref, rel=1e-12)


def test_cattaneo_force_golden(oracle):
    """examples/cattaneo.inp case 2 (T=1, F=2: Fx, Fy prescribed, Newton on the shift):
    examples/cattaneo.ref_out:88-139 -- shift 8.155e-3, Fx -3.218, C/A/S 177/45/132,
    ItCG 59, sensitivities dFx/dcksi -67.6 and dFy/dceta -117 as printed."""
    c = C.cattaneo_force_case()
    P = oracle.Problem()
    inp, h = C.to_oracle(oracle, c)
    assert P.calculate(inp, h) == 0
    fn, fx, fy, mz = P.forces()
    cksi, ceta, _ = P.creepages()
    el = P.eldiv()
    assert abs(fn - 9.195) < 5e-4 and abs(fx - (-3.218)) < 5e-4 and abs(fy) < 5e-4
    assert _sig(cksi, 8.155e-3) and _sig(ceta, -6.716e-9)
    assert (el > 0).sum() == 177 and (el == 1).sum() == 45 and (el == 2).sum() == 132
    assert P.iters()[3] == 59
    S = P.sens()
    assert float("%.3g" % S[1, 1]) == -67.6 and float("%.3g" % S[2, 2]) == -117.0


@pytest.mark.parametrize("row", C.CATTANEO_FORCE_TRACE, ids=lambda r: "nr%d" % r[0])
def test_cattaneo_force_newton_trace(oracle, row):
    """Each Newton iteration of the golden case (examples/cattaneo.ref_out:88-101), by
    stopping after `nr` iterations: creepages and forces to the printed 4 digits and the
    cumulative TangCG iterations exactly."""
    nr, cksi_r, fx_r, ceta_r, fy_r, itcg_r = row
    c = dict(C.cattaneo_force_case(), maxin=1)
    inp, h = C.to_oracle(oracle, c)
    inp.maxnr = nr
    P = oracle.Problem()
    assert P.calculate(inp, h) == 0
    fn, fx, fy, _ = P.forces()
    cksi, ceta, _ = P.creepages()
    assert _sig(cksi, cksi_r) and _sig(fx / (0.4 * fn), fx_r)
    assert _sig(ceta, ceta_r) and _sig(fy / (0.4 * fn), fy_r)
    assert P.iters()[3] == itcg_r


def test_gdsteady_stagnation_falls_back(oracle):
    """A batch patch on which GDsteady stagnates: SteadyGS takes over (src/m_solvpt.f90:466-485)
    and the result no longer depends on rounding-level changes of the input."""
    from contact_amd import workloads as W
    c = W.hertz_batch(n=2048)[381]
    res = []
    for f in (1.0, 1.0 + 1e-14):
        cc = dict(c)
        cc["fn"] = c["fn"] * f
        P = oracle.Problem()
        inp, h = C.to_oracle(oracle, cc)
        assert P.calculate(inp, h) == 0
        assert P.stagnated() == 3
        res.append(np.array(P.forces()[:3]))
    assert np.max(np.abs(res[0] - res[1])) <= 1e-7 * np.abs(res[0][0])


def test_force_prescribed_rolling_reaches_target(oracle):
    """T=3 with Fx prescribed (F=1, GDsteady) and with Fx, Fy prescribed (F=2, SteadyGS): the
    Newton iteration on the creepages reaches the prescribed forces within eps, and a
    solve with the creepages it found, prescribed (F=0), gives the same forces."""
    for force, gausei in ((1, 5), (2, 0)):
        c = C.hertz_case(6.0, 4.0, 24, 24, 8200.0, cphi=0.0005, eps=1e-9, maxgs=2000, maxin=40,
                         gausei=gausei)
        c["ceta"] = 0.0002
        c.update(force=force, fxrel=-0.6, fyrel=0.3)
        P = oracle.Problem()
        inp, h = C.to_oracle(oracle, c)
        assert P.calculate(inp, h) == 0
        fn, fx, fy, _ = P.forces()
        cksi, ceta, _ = P.creepages()
        assert abs(fx / (0.3 * fn) - (-0.6)) <= 2e-8
        if force == 2:
            assert abs(fy / (0.3 * fn) - 0.3) <= 2e-8
        else:
            assert ceta == 0.0002
        g = dict(c, force=0, cksi=cksi, ceta=ceta)
        Q = oracle.Problem()
        inp, h = C.to_oracle(oracle, g)
        assert Q.calculate(inp, h) == 0
        assert np.allclose(Q.forces()[1:3], (fx, fy), rtol=0, atol=1e-6 * fn)


# ---- Hertz option -2: curvature a1 and ellipticity aob (src/m_hertz.f90:95-213) -------------

@pytest.mark.parametrize("aob", [0.3, 1.0, 2.7])
def test_hertz_ellipticity_option(oracle, aob):
    """elliptic_to_curv finds the curvature b1 whose Hertz ellipse has aa/bb = aob.  Pinned by
    the other two Hertz options: ipotcn -1 with the b1 found gives the same ellipse, and
    ipotcn -3 with that ellipse's semi-axes gives back the curvatures a1 and b1."""
    from contact_amd import workloads as W
    a1 = 1.0 / 920.0
    base = dict(mx=10, my=10, scale=1.3, fn=1e5, tang=0)
    rc, _, P2 = oracle.solve_case(W.base_case(ipotcn=-2, a1=a1, aob=aob, **base))
    assert rc == 0
    h2 = P2.hertz()
    b1, aa, bb = h2[1], h2[2], h2[3]
    assert h2[0] == pytest.approx(a1, rel=1e-15)
    assert aa / bb == pytest.approx(aob, rel=1e-7)
    rc, _, P1 = oracle.solve_case(W.base_case(ipotcn=-1, a1=a1, b1=b1, **base))
    assert rc == 0
    h1 = P1.hertz()
    assert h1[2] == pytest.approx(aa, rel=1e-14) and h1[3] == pytest.approx(bb, rel=1e-14)
    rc, _, P3 = oracle.solve_case(W.base_case(ipotcn=-3, aa=aa, bb=bb, **base))
    assert rc == 0
    h3 = P3.hertz()
    assert h3[0] == pytest.approx(a1, rel=1e-7) and h3[1] == pytest.approx(b1, rel=1e-7)


def test_displacements_close_the_gap_and_the_shift(oracle):
    """The displacement differences us = A ps of a solution (soutpt, src/m_soutpt.f90:371-397)
    satisfy the contact conditions they were solved for: in contact the deformed gap
    h + un - pen is closed, and in adhesion under the Cattaneo shift (T=1, dq=1) the
    tangential displacement cancels the rigid shift: ux = -cksi, uy = 0."""
    c = C.cattaneo_case()
    c["eps"] = 1e-12
    P = oracle.Problem()
    inp, h = C.to_oracle(oracle, c)
    assert P.calculate(inp, h) == 0
    un, ux, uy = P.displacements()
    el = P.eldiv()
    hs, pen = P.h(), P.pen()
    con, adh = el > 0, el == 1
    assert adh.sum() > 20
    assert np.abs(un[con] + hs[con] - pen).max() <= 1e-12 * pen
    assert np.abs(ux[adh] + c["cksi"]).max() <= 1e-12 * c["cksi"]
    assert np.abs(uy[adh]).max() <= 1e-12 * c["cksi"]
    assert np.all(un[~con] == 0.0) and np.all(ux[~con] == 0.0)


# ---- transient rolling (T=2) pinned by its steady state --------------------------------------

def test_transient_rolling_reaches_the_steady_state(oracle):
    """Transient rolling (T=2, dq = dx) stepped with continuation (P=0: pv = ps, cv rolled by dq)
    and warm starts (I=3) at fixed creepages reaches a fixed point; its equations are then those
    of steady rolling (T=3), whose SteadyGS solution the table goldens pin.  So the fixed point
    must equal the T=3 solution: this pins set_prev_data, the rolled coefficients cv and csv,
    the pv terms of stang_rhs and TangCG's warm start (src/m_sdis.f90:208-317,
    src/m_visc.f90:319-358, src/m_stang.f90:749-951)."""
    base = C.hertz_case(6.0, 4.0, 30, 24, 80000.0, cksi=-0.0008, ceta=0.0003, eps=1e-12,
                        maxgs=2000, maxin=50, gausei=0)
    P3 = oracle.Problem()
    assert P3.calculate(*C.to_oracle(oracle, base)) == 0
    f3, p3, el3 = np.array(P3.forces()), P3.tractions(), P3.eldiv()
    P = oracle.Problem()
    dist = []
    for step in range(450):     # the distance to the fixed point falls by ~0.93 per step
        c = dict(base, tang=2, pvtime=2 if step == 0 else 0, iestim=0 if step == 0 else 3)
        assert P.calculate(*C.to_oracle(oracle, c)) == 0
        f = np.array(P.forces())
        dist.append(np.abs(f[:3] - f3[:3]).max() / abs(f3[1]))
    print("T=2 distance to the T=3 solution after 1, 100, 200, 450 steps:", dist[0], dist[99],
          dist[199], dist[-1])
    assert dist[0] > 1e-2 and dist[99] > 1e-5      # genuinely transient at first
    assert dist[-1] <= 1e-12
    p = P.tractions()
    assert max(np.abs(a - b).max() for a, b in zip(p, p3)) <= 1e-10 * np.abs(p3[0]).max()
    assert np.array_equal(P.eldiv(), el3)


def test_testbank_transient_chain_goldens(oracle):
    """testbank cases 13-16 (testbank/testbank.inp:245-301) against the printed results
    of testbank/testbank.ref_out: B=2 (prescribed Hertz pressure) shift with Fx prescribed, then
    transient rolling with continuation and I=1 at dq < dx (two steps) and at dq > dx, where the
    leading-edge correction (sxbnd/subnd) enters the right-hand side.  Pins the oracle's T=2 path,
    its force-prescribed Newton iteration and sensitivity, the Hertz option -2 and B=2."""
    P = oracle.Problem()
    for k, c in zip((13, 14, 15, 16), C.testbank_chain()):
        c = dict(c)
        if c["cksi"] is None:
            c["cksi"] = P.creepages()[0]
        assert P.calculate(*C.to_oracle(oracle, c)) == 0, k
        fn, fx, fy, mz = P.forces()
        el = P.eldiv()
        cksi = P.creepages()[0]
        g = C.TESTBANK_CHAIN_GOLDEN[k]
        got = (cksi, fx / (c["fstat"] * fn), int((el == 1).sum()), int((el == 2).sum()), P.iters()[3],
               P.sens()[1][1])
        print(k, got)
        assert abs(got[0] - g[0]) <= 0.6e-3 * abs(g[0]) + 1e-15, (k, got, g)   # 4 printed digits
        assert abs(got[1] - g[1]) <= 0.6e-4, (k, got, g)
        assert got[2:5] == g[2:5], (k, got, g)
        if g[5] is not None:
            assert abs(got[5] - g[5]) <= 0.06 if abs(g[5]) > 10 else abs(got[5] - g[5]) <= 0.006


def test_line_contact_hertz_options(oracle):
    """The 2D Hertz options are inverse to each other (hzcalc2d, src/m_hertz.f90:217-263): -5
    with half-length aa gives the curvature a1 for which -4 gives back aa; the prescribed
    semi-elliptical pressure (B=2) carries the normal force (its discrete integral to the
    grid's resolution) and is uniform in y over |y| <= bb."""
    from contact_amd import workloads as W
    base = dict(mx=40, my=20, bb=3.0, scale=1.2, fn=5e4, eps=1e-10, maxgs=500, maxin=30, tang=0)
    rc, _, P5 = oracle.solve_case(W.base_case(ipotcn=-5, aa=5.0, **base))
    assert rc == 0
    a1 = P5.hertz()[0]
    rc, _, P4 = oracle.solve_case(W.base_case(ipotcn=-4, a1=a1, **base))
    assert rc == 0
    assert P4.hertz()[2] == pytest.approx(5.0, rel=1e-14)
    pn = P5.tractions()[0].reshape(20, 40)
    xc1, yc1, dx, dy = P5.grid()
    assert pn.sum() * dx * dy == pytest.approx(5e4, rel=2e-2)
    inside = np.abs(yc1 + np.arange(20) * dy) <= 3.0
    assert np.allclose(pn[inside], pn[inside][0], rtol=0, atol=0) and np.all(pn[~inside] == 0)


def test_convexgs_carter_golden(oracle):
    """testbank case 6 (ConvexGS, steady rolling with dq = 5 dx and the leading-edge
    correction): every printed figure of testbank.ref_out, and Fy below the reference's print
    filter 0.5 eps mu Fn (src/m_soutpt.f90:536)."""
    c = C.carter_convexgs_case()
    P = oracle.Problem()
    assert P.calculate(*C.to_oracle(oracle, c)) == 0
    fn, fx, fy, mz = P.forces()
    el = P.eldiv()
    its = P.iters()
    g = C.CARTER_GOLDEN
    print("carter: fx", fx, "A/S", (el == 1).sum(), (el == 2).sum(), "iters", its, "pen", P.pen())
    assert round(fx) == g[0] and round(fx / (0.3 * fn), 4) == g[1]
    assert ((el > 0).sum(), (el == 1).sum(), (el == 2).sum()) == g[2:5]
    assert its[2] == g[5] and its[3] == g[6]
    assert float("%.3e" % P.pen()) == g[7] and round(P.tractions()[0].max(), 1) == g[8]
    assert abs(fy) < 0.5 * c["eps"] * 0.3 * fn


def test_convexgs_equals_other_solvers(oracle):
    """ConvexGS solves the same equations as TangCG (shift, T=1) and as SteadyGS (steady
    rolling at dq = dx, where A(cs - cv) p = A(cs) dp): at eps 1e-11 the converged forces agree
    to 1e-8 of Fn and the element states are identical."""
    for tang, other in ((1, 0), (3, 0)):
        base = C.hertz_case(5.0, 4.0, 22, 18, 6e4, cksi=-0.0009, ceta=0.0003, cphi=0.1e-3,
                            eps=1e-11, maxgs=3000, maxin=30, tang=tang, gausei=other)
        cg = dict(base, gausei=2, omegah=0.6, omegas=0.6)
        Pa, Pb = oracle.Problem(), oracle.Problem()
        assert Pa.calculate(*C.to_oracle(oracle, base)) == 0
        assert Pb.calculate(*C.to_oracle(oracle, cg)) == 0
        fa, fb = np.array(Pa.forces()[:3]), np.array(Pb.forces()[:3])
        print("T=%d" % tang, "G=%d" % other, fa, "ConvexGS", fb, Pb.iters())
        assert np.abs(fa - fb).max() <= 1e-8 * fa[0]
        assert np.array_equal(Pa.eldiv(), Pb.eldiv())


def test_testbank_panagiotopoulos_golden(oracle):
    """testbank case 9 (dissimilar materials, K = 0.4: the normal and tangential problems
    coupled through the cross influence coefficients and alternated by panprc,
    src/m_scontc.f90:356-553): every figure testbank.ref_out:951-969 prints for it."""
    c = C.panag_case()
    P = oracle.Problem()
    assert P.calculate(*C.to_oracle(oracle, c)) == 0
    g = C.PANAG_GOLDEN
    fn, fx, fy, mz = P.forces()
    assert round(fn, 3) == g[0] and round(fy, 4) == g[1] and round(mz, 3) == g[2]
    assert round(fx, 4) == round(-0.4718 * 0.6 * 2.468, 4)
    cksi = P.creepages()[0]
    assert float("%.3e" % cksi) == g[3] and round(fy / (0.6 * fn), 4) == g[4]
    assert float("%.3e" % P.pen()) == g[5]
    assert float("%.3e" % P.tractions()[0].max()) == g[6]
    el = P.eldiv()
    assert ((el > 0).sum(), (el == 1).sum(), (el == 2).sum()) == g[7:10]
    assert P.iters()[0] == P.iters()[1] == g[10]
    S = P.sens()
    assert round(S[1, 1]) == g[11] and round(S[2, 1], 1) == g[12]
