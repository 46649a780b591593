"""Case definitions shared by the parity tests, the smoke test and the benchmark.

Each case is a dict of cntc_* settings; `apply_cntc` issues them through the library and
`to_oracle` builds the equivalent input record of the CPU restatement.
"""
import math

import numpy as np

def hertz_case(aa, bb, mx, my, fn, cksi=0.0, ceta=0.0, cphi=0.0, fstat=0.3, eps=1e-6,
               maxgs=299, maxin=1, scale=None, tang=3, g=82000.0, nu=0.28, precision=0,
               gausei=5):
    """Hertzian patch with prescribed semi-axes (cntc_setHertzContact ipotcn=-3)."""
    if scale is None:
        scale = float(np.float32(mx) / np.float32(max(1, mx - 4)))
    return dict(tang=tang, norm=1, fn=fn, ipotcn=-3, mx=mx, my=my, aa=aa, bb=bb, scale=scale,
                cksi=cksi, ceta=ceta, cphi=cphi, fstat=fstat, g=g, nu=nu, eps=eps, maxgs=maxgs,
                maxin=maxin, veloc=10000.0, precision=precision, gausei=gausei)


def grid_case(mx, my, xl, yl, dx, dy, prmudf=None, h=None, norm=0, pen=0.0, fn=0.0, tang=1,
              cksi=0.0, ceta=0.0, cphi=0.0, fstat=0.3, eps=1e-8, maxgs=100, maxin=100,
              g=82000.0, nu=0.28, precision=0, dq=None):
    """Grid given by corner and steps (cntc_setPotContact ipotcn=1), gap quadratic or element-wise."""
    return dict(tang=tang, norm=norm, pen=pen, fn=fn, ipotcn=1, mx=mx, my=my, xl=xl, yl=yl,
                dx=dx, dy=dy, prmudf=prmudf, h=h, cksi=cksi, ceta=ceta, cphi=cphi, fstat=fstat,
                g=g, nu=nu, eps=eps, maxgs=maxgs, maxin=maxin, veloc=10000.0,
                precision=precision, dq=dq)


def cattaneo_case():
    """testbank.inp case 1: Cattaneo shift, penetration and creepage prescribed."""
    return grid_case(17, 17, -1.13333, -1.13333, .13333, .13333,
                     prmudf=[.0056566, 0.0, .0056566, 0.0, 0.0, 0.0], norm=0, pen=.01131,
                     tang=1, cksi=.00304, eps=1e-8, maxgs=100, maxin=100)


def cattaneo_force_case():
    """examples/cattaneo.inp case 2: T=1 shift with Fx = -7/8 mu Fn and Fy = 0 prescribed (F=2),
    Fn = 9.1954 (N=1), 19x19 grid, polyethylene (nu 0.42, G 200), mu 0.4, eps 1e-4."""
    c = grid_case(19, 19, -1.26667, -1.26667, .13333, .13333,
                  prmudf=[0.0100, 0.0, 0.0100, 0.0, 0.0, 0.0], norm=1, fn=9.1954, tang=1,
                  fstat=0.4, nu=0.42, g=200.0, eps=1e-4, maxgs=100, maxin=100)
    c.update(force=2, fxrel=-0.875, fyrel=0.0, gausei=0)
    return c


# examples/cattaneo.ref_out:88-102, the Newton iterations of case 2 as printed (4 digits):
# (iteration, cksi, Fx/(mu Fn), ceta, Fy/(mu Fn), cumulative TangCG iterations)
CATTANEO_FORCE_TRACE = [
    (1, 4.000e-06, -5.424e-04, 3.000e-06, -4.068e-04, 20),
    (2, 6.453e-03, -0.7410, 0.0, 5.200e-06, 33),
    (3, 7.620e-03, -0.8367, 4.460e-08, -4.188e-06, 42),
    (4, 8.087e-03, -0.8704, 8.681e-09, -2.461e-06, 50),
    (5, 8.150e-03, -0.8747, -1.243e-08, 6.659e-07, 55),
    (6, 8.155e-03, -0.8750, -6.716e-09, 9.615e-07, 59),
]


def table_cases(mx=11, ncase=3220):
    """The 3220 steady-rolling Hertz cases of src/test_table.f90:91-305 (B=0, T=3)."""
    ellip = [0.2, 0.5, 1.0, 2.0, 5.0]
    aa = [4.0, 6.0, 9.0, 12.0, 20.0]
    spin = [0.0, 0.5, 1.0, 2.0]
    vecangl = [-math.pi / 2, -math.pi / 3, -math.pi / 6, 0.0, math.pi / 6, math.pi / 3, math.pi / 2]
    vecmagn = [0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.7, 0.9, 1.1, 1.4, 1.7, 2.0, 2.5, 3.0, 3.5, 4.0,
               5.0, 6.0, 7.0, 8.0, 9.0, 10., 12.]
    fstat = 0.3
    out = []
    nv, na, ns = len(vecmagn), len(vecangl), len(spin)
    # rho, cp of the five ellipses (values printed by the reference at
    # testbank/test_table_mx11.ref_fx:1-5 to 4 digits; exact values from the closed form)
    rc = {}
    for icase in range(1, ncase + 1):
        ivmgn = (icase - 1) % nv + 1
        irem = (icase - ivmgn) // nv
        ivang = irem % na + 1
        irem = (irem - (ivang - 1)) // na
        ispin = irem % ns + 1
        irem = (irem - (ispin - 1)) // ns
        iell = irem % 5 + 1
        out.append(dict(icase=icase, iell=iell, aa=aa[iell - 1], bb=aa[iell - 1] / ellip[iell - 1],
                        cmgn=vecmagn[ivmgn - 1], cang=vecangl[ivang - 1], spin=spin[ispin - 1],
                        fstat=fstat, mx=mx))
    return out, rc


def table_golden():
    """Rows of testbank/test_table_mx11.ref_fx (tests/golden/test_table_mx11_fx.csv):
    Fx/(mu Fn), Fy/(mu Fn), Mz/(mu Fn cp), % slip per case."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "test_table_mx11_fx.csv")
    return np.loadtxt(path, delimiter=",", comments="#")


def table_case_inputs(O, r, gausei):
    """The cntc settings of table row r (src/test_table.f90:196-292): creepages scaled by
    mu cp / rho of the row's Hertz ellipse (rho, cp from the oracle's Hertz solution)."""
    base = hertz_case(r["aa"], r["bb"], r["mx"], r["mx"], 82000.0, fstat=0.3, eps=1e-6,
                      maxgs=299, maxin=1, gausei=gausei)
    P = O.Problem()
    P.calculate(to_oracle(O, base)[0])
    _, _, _, _, rho, cp = P.hertz()
    cksi = r["cmgn"] * 0.3 * math.cos(r["cang"]) * cp / rho
    ceta = r["cmgn"] * 0.3 * math.sin(r["cang"]) * cp / rho
    cphi = 0.3 * r["spin"] / rho
    c = hertz_case(r["aa"], r["bb"], r["mx"], r["mx"], 82000.0, cksi=cksi, ceta=ceta, cphi=cphi,
                   eps=1e-6, maxgs=299, maxin=1, gausei=gausei)
    return c, cp


def apply_cntc(cntc, ire, icp, c):
    """Issue the cntc_* set-calls of case `c` for problem (ire, icp)."""
    from contact_amd.workloads import apply_case
    apply_case(cntc, ire, icp, c)


def to_oracle(O, c):
    """The oracle input record of case `c` (same settings as apply_cntc)."""
    return O.case_input(c)


def testbank_chain():
    """testbank/testbank.inp:245-301, cases 13-16, as case dicts (mm, N): a Cattaneo shift with
    the Hertzian pressure prescribed (B=2) and Fx prescribed (F=1), then transient rolling
    (T=2, continuation P=0, regularised warm start I=1) at dq = 0.2 < dx = 0.2343 with Fx
    prescribed (two steps), then at dq = 0.8 > dx with the creepage prescribed and the Hertz
    geometry given by curvature and ellipticity (ipotcn -2): the leading-edge correction of
    src/m_leadedge.f90 is active only in the last case.  Between cases the creepage carries
    over (the .inp cases 14 and 15 give Fx, not cksi)."""
    from contact_amd import workloads as W
    common = dict(norm=1, fn=0.7563, fstat=0.4013, g=1.0, nu=0.28, eps=1e-6, maxgs=80, maxin=100,
                  gausei=0, mx=41, my=41, scale=1.17143, ceta=0.0, cphi=0.0)
    c13 = W.base_case(**common, tang=1, pvtime=2, iestim=0, bound=2, force=1, fxrel=0.657,
                      ipotcn=-3, aa=4.1, bb=4.1, cksi=0.0, veloc=1.0)
    c14 = W.base_case(**common, tang=2, pvtime=0, iestim=1, bound=0, force=1, fxrel=0.657,
                      ipotcn=-3, aa=4.1, bb=4.1, cksi=None, dq=0.2, veloc=1000.0)
    c16 = W.base_case(**common, tang=2, pvtime=0, iestim=1, bound=0, force=0, ipotcn=-2,
                      a1=0.2963e-2, aob=1.0, cksi=-0.002124, dq=0.8, veloc=1000.0)
    return [c13, c14, dict(c14), c16]


# testbank/testbank.ref_out: per case (creepage cksi (shift for case 13), Fx/(mu Fn), #adhesion,
# #slip, cumulative TangCG iterations, dFx/dcksi or None) as printed
TESTBANK_CHAIN_GOLDEN = {
    13: (-2.432e-02, 0.6570, 485, 488, 74, -22.4),      # :1296, :1330-1339
    14: (-2.175e-03, 0.6570, 668, 305, 77, -5.27),      # :1407, :1443-1450
    15: (-2.185e-03, 0.6570, 652, 321, 22, -5.21),      # :1514, :1536-1543
    16: (-2.124e-03, 0.6563, 610, 363, 15, None),       # :1605, :1642
}


def carter_convexgs_case():
    """testbank/testbank.inp:88-103, case 6: the Carter 2-D example (one row, dy = 200) in
    steady rolling with ConvexGS (G=2, omegah = omegas = 0.8) and dq = 5 dx; Fn prescribed,
    h = 0.001 x^2."""
    from contact_amd import workloads as W
    dx, dy = 0.04, 200.0
    return W.base_case(tang=3, norm=1, fn=35780.0, cksi=-0.00024, ceta=0.0, cphi=0.0, fstat=0.3,
                       g=82000.0, nu=0.28, eps=1e-5, maxgs=999, maxin=100, gausei=2, omegah=0.8,
                       omegas=0.8, ipotcn=3, mx=55, my=1, xl=-1.0450 - 0.5 * dx, yl=-0.5 * dy,
                       dx=dx, dy=dy, prmudf=[0.001, 0.0, 0.0, 0.0, 0.0, 0.0], dq=0.2,
                       veloc=30000.0, pvtime=2)


# testbank/testbank.ref_out:551-601 (case 6): Fx, Fx/(mu Fn), contact, adhesion, slip,
# NormCG iterations, ConvexGS sweeps, approach, pmax as printed
CARTER_GOLDEN = (6974.0, 0.6497, 51, 30, 21, 11, 163, 6.492e-3, 113.9)


def panag_case():
    """testbank/testbank.inp:187-203, case 9: dissimilar materials (poiss 0.1 / 0.28, gg 1 /
    81000, so K = 0.4) in steady rolling with SteadyGS, Fn and Fx/(mu Fn) = -0.4718 prescribed,
    h = 0.3332e-3 (x^2 + y^2) on a 25 x 25 grid: the Panagiotopoulos alternation of the normal
    and tangential problems (src/m_scontc.f90:356-553, up to 10 outer iterations)."""
    from contact_amd import workloads as W
    return W.base_case(tang=3, norm=1, fn=2.468, cksi=0.0, ceta=0.0, cphi=1.0e-4, fstat=0.6,
                       g=1.0, g2=0.81e5, nu=0.1, nu2=0.28, eps=1e-6, maxgs=80, maxin=100,
                       gausei=0, ipotcn=1, mx=25, my=25, xl=-13.75, yl=-13.75, dx=1.1, dy=1.1,
                       prmudf=[0.3332e-3, 0.0, 0.3332e-3, 0.0, 0.0, 0.0], dq=1.1, veloc=1000.0,
                       pvtime=2, force=1, fxrel=-0.4718, maxout=10)


# testbank/testbank.ref_out:951-963, 965-969 (case 9): Fn, Fy, Mz, cksi, Fy/(mu Fn), approach,
# pmax, contact, adhesion, slip, outer iterations, dFx/dcksi, dFy/dcksi as printed
PANAG_GOLDEN = (2.468, -0.2610, -1.470, 2.827e-4, -0.1762, 7.607e-2, 1.199e-2, 292, 140, 152,
                10, -272.0, 30.3)
