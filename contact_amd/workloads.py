"""Synthetic contact-patch workloads and a helper that issues them through the cntc_* calls.

A case is a plain dict of cntc_* settings (units: mm, N, N/mm^2, rad/mm).  `apply_case`
issues the set-calls for one (ire, icp) problem.  Generators:

* `hertz_40`       -- 40 x 40 Hertzian patch, steel on steel, Fn = 100 kN, steady rolling.
* `rough_patch`    -- N x N grid over 1.3x the Hertz semi-axes of the same wheel/rail
                      curvatures, gap = Hertz quadratic + seeded Gaussian roughness
                      (RMS 1 um, exponential correlation length 4 elements), Fn = 120 kN.
* `hertz_batch`    -- patches with sides log-uniform in [20, 256], semi-axis ratio log-uniform
                      in [0.2, 5], Fn uniform 20-150 kN, creepages uniform +-0.005, spin
                      uniform +-0.5 rad/m, friction uniform 0.2-0.4, one seed.
"""
import math

import numpy as np
from scipy import special

G_STEEL = 82000.0
NU_STEEL = 0.28
# GDsteady parameters of the reference's steady-rolling inputs (perfc_test/tang_problm_1c.inp:12)
GD_PARAMS = dict(fdecay=1.0, betath=0.05, kdowfb=1, d_ifc=2.0, d_lin=-1.0, d_cns=1.0, d_slp=2.6,
                 pow_s=1.0)


def hertz_semiaxes(a1, b1, fn, g=G_STEEL, nu=NU_STEEL):
    """Semi-axes (a along x, b along y) of the Hertz ellipse for gap a1 x^2 + b1 y^2 and force fn."""
    e_star = g / (1.0 - nu)
    zbla = b1 <= a1
    ratio = b1 / a1 if zbla else a1 / b1

    def eli(k):                       # (E - B) / B with B = (E - mc K)/m, m = k^2
        m = k * k
        if m < 1e-16:
            return 1.0, math.pi / 2, math.pi / 2
        K, E = special.ellipk(m), special.ellipe(m)
        mc = 1.0 - m
        B = (E - mc * K) / m
        return (E - B) / B, K, E

    lo, hi = 0.0, 1.0 - 1e-15
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if eli(mid)[0] > ratio:
            lo = mid
        else:
            hi = mid
    k = 0.5 * (lo + hi)
    _, K, E = eli(k)
    gg = math.sqrt(max(1e-40, 1.0 - k * k))
    sg = math.sqrt(gg)
    rho = 2.0 / (a1 + b1)
    cp = ((3.0 * fn * rho * E) / (4.0 * math.pi * e_star * sg)) ** (1.0 / 3.0)
    if zbla:
        return cp * sg, cp / sg
    return cp / sg, cp * sg


def base_case(**kw):
    c = dict(tang=3, norm=1, fn=0.0, pen=0.0, fstat=0.3, g=G_STEEL, nu=NU_STEEL, eps=1e-8,
             maxgs=999, maxin=20, veloc=10000.0, cksi=0.0, ceta=0.0, cphi=0.0, precision=0)
    c.update(kw)
    return c


def hertz_40(eps=1e-8):
    """40 x 40 grid over 1.3x the Hertz semi-axes, h = x^2/920 + y^2/600, Fn = 100 kN, xi = -0.001."""
    a1, b1 = 1.0 / 920.0, 1.0 / 600.0
    a, b = hertz_semiaxes(a1, b1, 1e5)
    mx = my = 40
    dx, dy = 2.6 * a / mx, 2.6 * b / my
    return base_case(ipotcn=1, mx=mx, my=my, xl=-1.3 * a, yl=-1.3 * b, dx=dx, dy=dy,
                     prmudf=[a1, 0.0, b1, 0.0, 0.0, 0.0], fn=1e5, cksi=-0.001, eps=eps)


def roughness(mx, my, rms, corr_len, seed):
    """Seeded Gaussian random field with exponential correlation (length in elements)."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((my, mx))
    kx = np.fft.fftfreq(mx) * 2 * np.pi
    ky = np.fft.fftfreq(my) * 2 * np.pi
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    psd = (1.0 + k2 * corr_len ** 2) ** -1.5      # 2-D spectrum of exp(-r / corr_len)
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * np.sqrt(psd)))
    f -= f.mean()
    return f * (rms / f.std())


def rough_patch(n=512, seed=1234, fn=1.2e5, eps=1e-8, precision=0):
    """Rough non-Hertzian patch on an n x n grid (element-wise gap, ibase=9)."""
    a1, b1 = 1.0 / 920.0, 1.0 / 600.0
    a, b = hertz_semiaxes(a1, b1, fn)
    dx, dy = 2.6 * a / n, 2.6 * b / n
    xl, yl = -1.3 * a, -1.3 * b
    x = xl + (np.arange(n) + 0.5) * dx
    y = yl + (np.arange(n) + 0.5) * dy
    h = a1 * x[None, :] ** 2 + b1 * y[:, None] ** 2 + roughness(n, n, 1e-3, 4.0, seed)
    return base_case(ipotcn=1, mx=n, my=n, xl=xl, yl=yl, dx=dx, dy=dy, h=h.ravel(), fn=fn,
                     cksi=0.002, ceta=0.001, cphi=0.1e-3, eps=eps, precision=precision)


def hertz_batch(n=2048, seed=2024, eps=1e-8, precision=0, side_range=(20, 256)):
    """Batch of Hertzian steady-rolling patches with randomised sizes and loads."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(side_range[0]), np.log(side_range[1])
    cases = []
    for _ in range(n):
        mx = int(round(np.exp(rng.uniform(lo, hi))))
        my = int(round(np.exp(rng.uniform(lo, hi))))
        ratio = float(np.exp(rng.uniform(np.log(0.2), np.log(5.0))))
        fn = float(rng.uniform(2e4, 1.5e5))
        cksi = float(rng.uniform(-0.005, 0.005))
        ceta = float(rng.uniform(-0.005, 0.005))
        cphi = float(rng.uniform(-0.5, 0.5)) * 1e-3
        mu = float(rng.uniform(0.2, 0.4))
        c = 6.24 * (fn / 1e5) ** (1.0 / 3.0)   # sqrt(a b) of steel wheel/rail contact at fn
        aa, bb = c * math.sqrt(ratio), c / math.sqrt(ratio)
        cases.append(base_case(ipotcn=-3, mx=mx, my=my, aa=aa, bb=bb, scale=1.3, fn=fn,
                               cksi=cksi, ceta=ceta, cphi=cphi, fstat=mu, eps=eps,
                               precision=precision))
    return cases


def apply_case(cntc, ire, icp, c):
    """Issue the cntc_* set-calls of case `c` for problem (ire, icp)."""
    cntc.setflags(ire, icp, [1972, 1973, 2009], [c["tang"], c["norm"], c.get("precision", 0)])
    if c.get("pvtime") is not None:
        cntc.setflags(ire, icp, [1970, 1983], [c["pvtime"], c.get("iestim", 0)])
    # poiss 1, 2 and gg 1, 2: "nu2" and "g2" set the second body of dissimilar materials (K != 0)
    cntc.setmaterialparameters(ire, icp, 0, [c["nu"], c.get("nu2", c["nu"]), c["g"], c.get("g2", c["g"])])
    cntc.setfrictionmethod(ire, icp, 0, [c["fstat"], c["fstat"]])
    cntc.setreferencevelocity(ire, icp, c.get("veloc", 10000.0))
    if c.get("gausei", 5) == 5:
        g = GD_PARAMS
        cntc.setsolverflags(ire, icp, 5, [c["maxgs"], c["maxin"], 30, c.get("maxout", 1), g["kdowfb"]],
                            [c["eps"], g["fdecay"], g["betath"], g["d_ifc"], g["d_lin"],
                             g["d_cns"], g["d_slp"], g["pow_s"]])
    elif c["gausei"] in (2, 3):
        # G=2 ConvexGS, G=3 fixed relaxation: omegah, omegas (inislp, omgslp unused with L=0)
        cntc.setsolverflags(ire, icp, c["gausei"], [c["maxgs"], c["maxin"], 30, c.get("maxout", 1), 1],
                            [c["eps"], c.get("omegah", 0.5), c.get("omegas", 0.5), 0.9])
    else:
        # G=0: the default solvers (SteadyGS for T=3), src/test_table.f90:151-160
        cntc.setsolverflags(ire, icp, 0, [c["maxgs"], c["maxin"], 30, c.get("maxout", 1)], [c["eps"]])
    if c["ipotcn"] == -3:
        cntc.sethertzcontact(ire, icp, -3, [c["mx"], c["my"], c["aa"], c["bb"], c["scale"]])
    elif c["ipotcn"] == -2:
        cntc.sethertzcontact(ire, icp, -2, [c["mx"], c["my"], c["a1"], c["aob"], c["scale"]])
    elif c["ipotcn"] == -1:
        cntc.sethertzcontact(ire, icp, -1, [c["mx"], c["my"], c["a1"], c["b1"], c["scale"]])
    elif c["ipotcn"] in (-4, -5):
        first = c["a1"] if c["ipotcn"] == -4 else c["aa"]
        cntc.sethertzcontact(ire, icp, c["ipotcn"], [c["mx"], c["my"], first, c["bb"], c["scale"]])
    else:
        cntc.setpotcontact(ire, icp, 1, [c["mx"], c["my"], c["xl"], c["yl"], c["dx"], c["dy"]])
        if c.get("h") is not None:
            cntc.setundeformeddistc(ire, icp, 9, np.asarray(c["h"], dtype=np.float64))
        else:
            cntc.setundeformeddistc(ire, icp, 1, c["prmudf"])
    if c.get("dq") is not None:
        cntc.setrollingstepsize(ire, icp, 0.0, c["dq"])
    if c["norm"] == 1:
        cntc.setnormalforce(ire, icp, c["fn"])
    else:
        cntc.setpenetration(ire, icp, c["pen"])
    cntc.setcreepages(ire, icp, c["cksi"], c["ceta"], c["cphi"])
    if c.get("force") is not None:
        cntc.setflags(ire, icp, [1974], [c["force"]])
        cntc.settangentialforces(ire, icp, c.get("fxrel", 0.0), c.get("fyrel", 0.0))
    if c.get("rztang") is not None:
        cntc.setflags(ire, icp, [1980], [c["rztang"]])
    if c.get("bound") is not None:
        cntc.setflags(ire, icp, [1971], [c["bound"]])
    if c.get("exrhs") is not None:
        cntc.setextrarigidslip(ire, icp, c["exrhs"][0], c["exrhs"][1])
    if c.get("sens") is not None:
        cntc.setflags(ire, icp, [1988], [c["sens"]])
    if c.get("mxsens") is not None or c.get("epsens") is not None:
        cntc.setsolverflags(ire, icp, 6, [c.get("mxsens", 30)], [c.get("epsens", 1e-4)])
