"""Sensitivities through the C ABI (cntc_getsensitivities, CNTC_ic_sens = 2, 3) against the
oracle: dFn/dpen by one NormCG (src/m_solvpn.f90:531-650) and the creepage columns by three
tangential solves (src/m_solvpt.f90:544-796), with the reference's own sensitivity solver
settings (mxsens 30, epsens 1e-4) unless a case sets tighter ones.

Tolerance: each column within 1e-5 of its largest entry.  Both sides run the same iterations
in FP64; the difference quotients divide force changes of about 1e-4 relative by the step, so
rounding differences of the solves (1e-12 relative) reach about 1e-8 of the column.
"""
import numpy as np
import pytest

from tests import cases as C

pytestmark = pytest.mark.gpu

TOL = 1e-5


def gpu_sens(cntc, c, ire):
    C.apply_cntc(cntc, ire, 1, c)
    rc = cntc.calculate(ire, 1)
    return rc, cntc.getsensitivities(ire, 1), cntc.getcontactforces(ire, 1)


def oracle_sens(O, c):
    P = O.Problem()
    inp, h = C.to_oracle(O, c)
    rc = P.calculate(inp, h)
    return rc, P.sens(), P.forces()


def check(name, g, o, cols=range(4)):
    rc_g, S_g, f_g = g
    rc_o, S_o, f_o = o
    print(name, "rc", rc_g, rc_o, "\ngpu\n", S_g, "\noracle\n", S_o, flush=True)
    assert rc_g == rc_o == 0
    for k in cols:
        scale = max(np.abs(S_o[:, k]).max(), 1e-300)
        assert np.abs(S_g[:, k] - S_o[:, k]).max() <= TOL * scale, (name, k)


@pytest.mark.parametrize("gausei", [5, 0])
def test_sens_steady_rolling(cntc, oracle, gausei):
    """T=3 partial slip with spin, GDsteady (G=5) or SteadyGS (G=0), S=3."""
    c = C.hertz_case(6.0, 4.0, 30, 30, 8200.0, cksi=0.0004, ceta=0.0002, cphi=0.001, eps=1e-8,
                     maxgs=999, maxin=30, gausei=gausei)
    c["sens"] = 3
    check("rolling G=%d" % gausei, gpu_sens(cntc, c, 21 + gausei), oracle_sens(oracle, c))


def test_sens_shift_tangcg(cntc, oracle):
    """testbank case 1 (Cattaneo shift, T=1, TangCG, N=0 on 17x17, below 700 elements), S=3."""
    c = C.cattaneo_case()
    c["sens"] = 3
    check("cattaneo", gpu_sens(cntc, c, 31), oracle_sens(oracle, c))


def test_sens_transient_tangcg(cntc, oracle):
    """T=2 transient rolling from rest (TangCG), S=3."""
    c = C.hertz_case(6.0, 6.0, 24, 24, 8200.0, cksi=0.0003, eps=1e-8, maxgs=999, maxin=30, tang=2)
    c["sens"] = 3
    check("transient", gpu_sens(cntc, c, 32), oracle_sens(oracle, c))


@pytest.mark.parametrize("mx", [20, 40])
def test_sens_normal_only(cntc, oracle, mx):
    """S=2: dFn/dpen only (20x20 forms un after the clean-up, 40x40 reuses the expansion
    product); the creepage columns stay zero."""
    c = C.hertz_case(6.0, 6.0, mx, mx, 8200.0, cksi=0.0003, eps=1e-8, maxgs=999, maxin=30)
    c["sens"] = 2
    g = gpu_sens(cntc, c, 33)
    o = oracle_sens(oracle, c)
    check("normal %d" % mx, g, o, cols=[0])
    assert np.all(g[1][:, 1:] == 0.0) and np.all(g[1][1:, 0] == 0.0)


def test_sens_tight_tolerance(cntc, oracle):
    """Tight sensitivity solves (G=6: mxsens 2000, epsens 1e-12) agree more closely."""
    c = C.hertz_case(6.0, 6.0, 30, 30, 8200.0, cksi=0.0005, cphi=0.001, eps=1e-11, maxgs=3000,
                     maxin=40)
    c.update(sens=3, mxsens=2000, epsens=1e-12)
    g = gpu_sens(cntc, c, 34)
    o = oracle_sens(oracle, c)
    check("tight", g, o)
    for k in range(4):
        scale = np.abs(o[1][:, k]).max()
        assert np.abs(g[1][:, k] - o[1][:, k]).max() <= 1e-7 * scale


def test_sens_units_and_layout(cntc):
    """SI units (CNTC_if_units = CNTC_un_si: m, N, Pa) scale the inputs and outputs as
    src/contact_addon.f90:5967-5990: dFn/dpen per m, d/dcphi per rad/m, Mz in N m; entries
    beyond 4x4 are left as they were."""
    import ctypes
    c = C.hertz_case(6.0, 4.0, 24, 24, 8200.0, cksi=0.0004, cphi=0.001, eps=1e-8, maxgs=999,
                     maxin=30)
    c["sens"] = 3
    rc, S, _ = gpu_sens(cntc, c, 35)
    assert rc == 0
    si = dict(c, aa=c["aa"] * 1e-3, bb=c["bb"] * 1e-3, g=c["g"] * 1e6, cphi=c["cphi"] * 1e3)
    cntc.setflags(36, 1, [1933], [1936])
    C.apply_cntc(cntc, 36, 1, si)
    assert cntc.calculate(36, 1) == 0
    lenout, lenin = 5, 6
    v = np.full(lenout * lenin, 7.5)
    i = lambda x: ctypes.byref(ctypes.c_int(x))
    cntc.lib().cntc_getsensitivities(i(36), i(1), i(lenout), i(lenin),
                                     v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    M = v.reshape(lenin, lenout).T
    expect = np.diag([1.0, 1.0, 1.0, 1e-3]) @ S @ np.diag([1e3, 1.0, 1.0, 1e-3])
    for k in range(4):
        scale = np.abs(expect[:, k]).max()
        assert np.abs(M[:4, k] - expect[:, k]).max() <= 1e-8 * scale, k
    assert np.all(M[4, :] == 7.5) and np.all(M[:, 4:] == 7.5)


def test_sens_batch_equals_alone(cntc):
    """Sensitivity solves batched with other patches give the same bits as alone."""
    cases = []
    for i, (cx, ph) in enumerate([(0.0004, 0.0), (0.0002, 0.002), (0.0, 0.001), (0.0006, 0.0005)]):
        c = C.hertz_case(6.0, 4.0 + i, 20 + 4 * i, 20 + 4 * i, 8200.0, cksi=cx, cphi=ph, eps=1e-8,
                         maxgs=999, maxin=30, gausei=5 if i % 2 == 0 else 0)
        c["sens"] = 3 if i != 2 else 2
        cases.append(c)
    alone = []
    for i, c in enumerate(cases):
        rc, S, _ = gpu_sens(cntc, c, 40 + i)
        assert rc == 0
        alone.append(S)
    for i, c in enumerate(cases):
        C.apply_cntc(cntc, 40 + i, 1, c)
    err = cntc.calculatebatch([40 + i for i in range(len(cases))], [1] * len(cases))
    assert list(err) == [0] * len(cases)
    for i in range(len(cases)):
        assert np.array_equal(cntc.getsensitivities(40 + i, 1), alone[i]), i
