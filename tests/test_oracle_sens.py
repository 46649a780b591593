"""CPU checks of the oracle's sensitivities (sens_norm, src/m_solvpn.f90:531-650; sens_tang and
char_creep, src/m_solvpt.f90:497-796; linrol, src/m_hertz.f90:1147-1247).

The reference's outputs hold no sensitivity computed this way (the DF/DKSI columns of its
example outputs come from the F=1 force iteration, which is outside this path), so the oracle
is pinned by:
  - Kalker's linear theory: at zero creepage the computed d(Fx/(mu Fn))/dcksi equals
    -G c^2 C11 / (mu Fn) = -1 / ch_ksi, and likewise for ceta (C22) and cphi (C23);
  - Hertz: dFn/dpen = 1.5 Fn / (pen - hmin);
  - brute force: two complete solves at the perturbed and unperturbed inputs give the same
    difference quotient when the sensitivity solve runs to a tight tolerance;
  - the table's nodes and the continuity of linrol's asymptotic branches.
"""
import math

import numpy as np
import pytest

from tests import cases as C


def _case(tang, mx=30, aa=6.0, bb=6.0, **kw):
    c = C.hertz_case(aa, bb, mx, mx, 8200.0, tang=tang, eps=1e-12, maxgs=5000, maxin=50, **kw)
    c["sens"] = 3
    c["epsens"] = 1e-13
    c["mxsens"] = 5000
    return c


def _solve(oracle, c):
    rc, f, P = oracle.solve_case(c)
    assert rc == 0, rc
    fn, fx, fy, mz = f
    mu = c["fstat"]
    return np.array([fn, fx / (mu * fn), fy / (mu * fn), mz]), P


def _char_creep(oracle, P, c):
    """src/m_solvpt.f90:497-540 from the oracle's element division and forces."""
    el = P.eldiv()
    mx, my = P.shape
    _, _, dx, dy = P.grid()
    inn = el.reshape(my, mx) >= 1
    ncon = int(inn.sum())
    cc = math.sqrt(ncon * dx * dy / math.pi)
    iy, ix = np.nonzero(inn)
    a = max(1, ix.max() - ix.min() + 1) * dx * 0.5
    b = max(1, iy.max() - iy.min() + 1) * dy * 0.5
    c11, c22, c23 = oracle.linrol(c["nu"], a / b)
    fn = P.forces()[0]
    k = c["fstat"] * fn / c["g"]
    return k / (cc ** 2 * c11), k / (cc ** 2 * c22), k / (cc ** 3 * c23)


# ---- linrol ------------------------------------------------------------------------------------

def test_linrol_kalker_circle(oracle):
    """a/b = 1, nu = 1/4: C11 = 4.12, C22 = 3.67, C23 = 1.47 (Kalker's table, as reproduced in
    K.L. Johnson, Contact Mechanics, table 8.1); both branches agree at a/b = 1."""
    c = oracle.linrol(0.25, 1.0)
    assert c == pytest.approx((4.12, 3.67, 1.47), abs=1e-6)
    assert oracle.linrol(0.25, 1.0 - 1e-12) == pytest.approx(oracle.linrol(0.25, 1.0 + 1e-12), rel=1e-9)
    # nu = 0: C11 = C22 (no Poisson coupling)
    c0 = oracle.linrol(0.0, 1.0)
    assert c0[0] == pytest.approx(c0[1], abs=1e-6)


@pytest.mark.parametrize("nu", [0.0, 0.25, 0.28, 0.5])
def test_linrol_asymptotes_continuous(oracle, nu):
    """The slender-ellipse formulas (g < 0.101) join the table at g = 0.1 within a few percent,
    for both a/b < 1 and a/b > 1: a dropped term or a wrong sign there breaks this."""
    for lo, hi in ((0.1009, 0.1011), (1 / 0.1009, 1 / 0.1011)):
        a, b = oracle.linrol(nu, lo), oracle.linrol(nu, hi)
        for x, y in zip(a, b):
            assert abs(x - y) <= 0.08 * abs(y), (nu, lo, a, b)


def test_linrol_monotone_in_nu_and_g(oracle):
    """Lagrange interpolation of 1/C in nu reproduces the three nu nodes and stays between them;
    C11 grows with nu and the coefficients grow towards slender ellipses along the rolling
    direction (a/b -> 0 is case 0's asymptote pi^2/4/(1-nu))."""
    for g in (0.15, 0.45, 0.8):
        c = [oracle.linrol(nu, g)[0] for nu in (0.0, 0.1, 0.25, 0.4, 0.5)]
        assert all(c[i] < c[i + 1] for i in range(4))
    assert oracle.linrol(0.25, 0.01)[0] == pytest.approx(math.pi ** 2 / 4 / 0.75, rel=1e-12)
    c1 = [oracle.linrol(0.25, 1.0 / g)[0] for g in (1.0, 0.7, 0.4, 0.2)]
    assert all(c1[i] < c1[i + 1] for i in range(3))


# ---- sens_norm -----------------------------------------------------------------------------------

def test_sens_norm_hertz(oracle):
    """dFn/dpen of a Hertzian contact: 1.5 Fn / (pen - hmin), within the grid error."""
    c = _case(0, mx=40)
    c["sens"] = 2
    f, P = _solve(oracle, c)
    S = P.sens()
    pen, hmin = P.pen(), P.h().min()
    assert S[0, 0] == pytest.approx(1.5 * f[0] / (pen - hmin), rel=5e-4)
    assert np.all(S[1:, 0] == 0.0) and np.all(S[:, 1:] == 0.0)


@pytest.mark.parametrize("mx", [20, 30])
def test_sens_norm_brute_force(oracle, mx):
    """One NormCG at pen + dpen from the converged state equals a complete solve there
    (dpen = 0.0067 max(un) = 0.0067 (pen - hmin) at convergence); 20x20 takes the path below
    700 elements, where un is formed after the clean-up of the pressures.  The quadratic gap of
    a Hertz solution is fixed on its grid, so that only pen changes."""
    _, H = _solve(oracle, dict(_case(0, mx=mx), sens=1))
    a1, b1 = H.hertz()[:2]
    xc1, yc1, dx, dy = H.grid()
    pen = H.pen()
    base = C.grid_case(mx, mx, xc1 - 0.5 * dx, yc1 - 0.5 * dy, dx, dy,
                       prmudf=[a1, 0.0, b1, 0.0, 0.0, 0.0], norm=0, pen=pen, tang=0,
                       eps=1e-12, maxgs=5000, maxin=50)
    f0, P = _solve(oracle, dict(base, sens=2, epsens=1e-13, mxsens=5000))
    hmin = P.h().min()
    dpen = 0.0067 * (pen - hmin)
    f1, _ = _solve(oracle, dict(base, pen=pen + dpen))
    fd = (f1[0] - f0[0]) / dpen
    assert P.sens()[0, 0] == pytest.approx(fd, rel=1e-7)
    assert fd == pytest.approx(1.5 * f0[0] / (pen - hmin), rel=5e-3)


# ---- sens_tang -------------------------------------------------------------------------------------

@pytest.mark.parametrize("ab", [(6.0, 6.0), (6.0, 3.0), (3.0, 6.0)])
def test_sens_tang_linear_theory(oracle, ab):
    """Steady rolling at zero creepage adheres everywhere and its sensitivities are Kalker's
    linear theory: d(Fx/(mu Fn))/dcksi = -1/ch_ksi, d(Fy/(mu Fn))/dceta = -1/ch_eta and
    d(Fy/(mu Fn))/dcphi = -1/ch_phi, up to the discretisation error of a 30x30 grid."""
    c = _case(3, aa=ab[0], bb=ab[1])
    f, P = _solve(oracle, c)
    S = P.sens()
    ch = _char_creep(oracle, P, c)
    assert S[1, 1] * ch[0] == pytest.approx(-1.0, abs=0.025)
    assert S[2, 2] * ch[1] == pytest.approx(-1.0, abs=0.025)
    assert S[2, 3] * ch[2] == pytest.approx(-1.0, abs=0.05)


@pytest.mark.parametrize("tang", [1, 3])
def test_sens_tang_brute_force(oracle, tang):
    """Partial slip: each column equals the difference quotient of two complete solves with the
    creepage raised by 1e-4 ch.  In rolling the spin's rigid slip is taken about x + dq/6
    (src/m_sdis.f90:498-583) while sens_tang uses x, so the spin column is compared with a
    solve that also lowers ceta by dq/6 * dcphi."""
    c = _case(tang, mx=30 if tang == 3 else 16, cksi=0.0005, cphi=0.001 if tang == 3 else 0.0)
    f0, P = _solve(oracle, c)
    S = P.sens()
    ch = _char_creep(oracle, P, c)
    _, _, dx, _ = P.grid()
    xofs = dx / 6.0 if tang == 3 else 0.0
    for k, name in enumerate(("cksi", "ceta", "cphi")):
        d = 1e-4 * ch[k]
        c2 = dict(c, sens=1)
        c2[name] = c[name] + d
        if name == "cphi":
            c2["ceta"] = c["ceta"] - xofs * d
        f1, _ = _solve(oracle, c2)
        fd = (f1 - f0) / d
        scale = np.abs(S[1:, 1 + k]).max()
        assert np.abs(fd[1:] - S[1:, 1 + k]).max() <= 1e-6 * scale, (name, fd, S[:, 1 + k])


def test_sens_tang_brute_force_has_teeth(oracle):
    """The comparison above resolves the perturbation size: twice the step of char_creep
    changes the partial-slip moment quotient by far more than its tolerance."""
    c = _case(3, cksi=0.0005, cphi=0.001)
    f0, P = _solve(oracle, c)
    S = P.sens()
    ch = _char_creep(oracle, P, c)
    d = 2e-4 * ch[0]
    f1, _ = _solve(oracle, dict(c, sens=1, cksi=c["cksi"] + d))
    fd = (f1 - f0) / d
    assert np.abs(fd[1:] - S[1:, 1]).max() > 1e-5 * np.abs(S[1:, 1]).max()


def test_sens_default_and_reset(oracle):
    """S2 = 1 (the default) computes nothing; I=0 clears what an earlier case left."""
    c = _case(3, cksi=0.0005)
    P = oracle.Problem()
    inp, _ = C.to_oracle(oracle, c)
    assert P.calculate(inp) == 0
    assert np.count_nonzero(P.sens()) >= 9
    inp, _ = C.to_oracle(oracle, dict(c, sens=1))
    assert P.calculate(inp) == 0
    assert np.all(P.sens() == 0.0)
