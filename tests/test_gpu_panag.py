"""Dissimilar materials (K != 0) through the cntc_* C ABI: the normal and tangential problems
coupled through the cross influence coefficients and alternated by the Panagiotopoulos loop
(src/m_scontc.f90:356-553), with the coupling terms in the normal gap (src/m_snorm.f90:112-119)
and in the tangential right-hand side (src/m_stang.f90:859-864).

testbank case 9 (eps 1e-6, as the reference runs it) is checked against every figure that
testbank.ref_out prints for it.  At eps 1e-10 the forces agree with the oracle to 1e-8 of Fn,
the tractions to 1e-7 of max pn and the element states are identical, for SteadyGS, GDsteady
and a shift by TangCG, and for SteadyGS with FP32 products in the refinement or with every
product dense.
"""
import numpy as np
import pytest

from tests import cases as C
from tests.test_gpu_parity import compare, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def test_panagiotopoulos_golden(cntc, oracle):
    c = C.panag_case()
    g = run_gpu(cntc, c, ire=640)
    o = run_oracle(oracle, c)
    compare(g, o, 1e-5, 1e-4, "case 9")
    gold = C.PANAG_GOLDEN
    fn, fx, fy, mz = g[1]
    its = cntc.getiterations(640, 1)
    print("case 9 gpu: forces", g[1], "iterations", its, "pen", cntc.getpenetration(640, 1),
          flush=True)
    assert round(fn, 3) == gold[0] and round(fy, 4) == gold[1] and round(mz, 3) == gold[2]
    assert round(fy / (0.6 * fn), 4) == gold[4]
    assert float("%.3e" % cntc.getpenetration(640, 1)) == gold[5]
    assert float("%.3e" % g[2][0].max()) == gold[6]
    el = g[3]
    assert ((el > 0).sum(), (el == 1).sum(), (el == 2).sum()) == gold[7:10]
    assert its[0] == its[2] == gold[10]


CASES = {
    "steady SteadyGS": dict(C.panag_case(), eps=1e-10, maxgs=999),
    "steady GDsteady": dict(C.panag_case(), eps=1e-10, maxgs=999, gausei=5),
    "shift TangCG": dict(C.panag_case(), eps=1e-10, maxgs=999, tang=1, force=0, cksi=2.0e-4),
}


@pytest.mark.parametrize("name", list(CASES))
def test_coupled_vs_oracle(cntc, oracle, name):
    c = CASES[name]
    ire = 641 + list(CASES).index(name)
    g = run_gpu(cntc, c, ire=ire)
    o = run_oracle(oracle, c)
    compare(g, o, 1e-8, 1e-7, name)
    its = cntc.getiterations(ire, 1)
    print(name, "outer iterations gpu", its[0], "oracle", o[4].iters()[0], flush=True)
    assert its[0] > 1


def _sequence(base, tang, dq=None):
    """Three cases on one problem: no previous case (P=2), then two continuations (P=0)."""
    c = dict(base, tang=tang, force=0, eps=1e-10, maxgs=999)
    if dq is not None:
        c["dq"] = dq
    return [dict(c, pvtime=2, cksi=2.0e-4), dict(c, pvtime=0, cksi=3.0e-4),
            dict(c, pvtime=0, cksi=1.0e-4)]


K0 = dict(C.panag_case(), nu2=0.1, g2=1.0)
SEQUENCES = {
    "shift K=0": _sequence(K0, 1),
    "shift K=0.4": _sequence(C.panag_case(), 1),
    "transient K=0.4 dq=dx": _sequence(C.panag_case(), 2, dq=1.1),
}


@pytest.mark.parametrize("name", list(SEQUENCES))
def test_continuation_vs_oracle(cntc, oracle, name):
    """Shift and transient rolling continuing a previous case (P=0): the previous tractions'
    displacements A' pv enter the tangential right-hand side (src/m_stang.f90:855-873), with
    the normal-tangential blocks when K != 0.  Every case of the sequence matches the oracle
    run through the same sequence."""
    ire = 650 + list(SEQUENCES).index(name)
    P = oracle.Problem()
    for k, c in enumerate(SEQUENCES[name]):
        g = run_gpu(cntc, c, ire=ire)
        assert P.calculate(*C.to_oracle(oracle, c)) == 0
        o = (0, P.forces(), P.tractions(), P.eldiv(), P)
        compare(g, o, 1e-8, 1e-7, "%s, case %d" % (name, k))


def test_coupled_transient_ledge_after_previous_case_refused(cntc):
    """K != 0 in transient rolling at dq != dx continuing a previous case needs A'_tn pv_n
    on the interior but not at the leading edge, which is not built: the call returns
    CNTC_err_other (-99) without solving, and the same case with P=2 solves."""
    c = dict(C.panag_case(), tang=2, force=0, cksi=2.0e-4, dq=2.2, pvtime=0)
    assert run_gpu(cntc, c, ire=644)[0] == -99
    assert run_gpu(cntc, dict(c, pvtime=2), ire=645)[0] == 0


def test_coupled_fp32_refinement_and_dense(cntc, oracle, monkeypatch):
    """The coupled steady-rolling case with FP32 products inside the FP64 refinement
    (CNTC_if_precision 1), and with every product dense on FP64 MFMA (CNTC_DENSE=2, the
    cross blocks included): both match the oracle's FP64 solution to the same bar."""
    c = dict(C.panag_case(), eps=1e-10, maxgs=999)
    o = run_oracle(oracle, c)
    g = run_gpu(cntc, dict(c, precision=1), ire=646)
    compare(g, o, 1e-8, 1e-7, "coupled, FP32 products in the refinement")
    monkeypatch.setenv("CNTC_DENSE", "2")
    g = run_gpu(cntc, c, ire=647)
    monkeypatch.delenv("CNTC_DENSE")
    compare(g, o, 1e-8, 1e-7, "coupled, dense products")


def test_coupled_in_a_batch_equals_single(cntc, oracle):
    """Coupled (K != 0) and uncoupled patches solved together by cntc_calculatebatch: each
    coupled patch's forces and tractions are bitwise those of the same patch solved alone,
    and the batch's products of different shapes (normal from tangential, tangential from
    normal, all nine blocks) share launches with the uncoupled ones."""
    cs = [dict(C.panag_case(), eps=1e-10, maxgs=999, cphi=1e-4 * (k + 1)) for k in range(3)]
    cs += [C.hertz_case(4.0 + k, 6.0, 14 + 3 * k, 12, 5e4, cksi=-0.0005 * (k + 1), eps=1e-8,
                        maxgs=999, maxin=20) for k in range(3)]
    single = [run_gpu(cntc, c, ire=660 + k) for k, c in enumerate(cs)]
    for k, c in enumerate(cs):
        C.apply_cntc(cntc, 670 + k, 1, c)
    err = cntc.calculatebatch([670 + k for k in range(len(cs))], [1] * len(cs))
    assert (err == 0).all()
    for k in range(len(cs)):
        f = cntc.getcontactforces(670 + k, 1)
        pn, px, py = cntc.gettractions(670 + k, 1)
        assert f == single[k][1], k
        assert np.array_equal(pn.ravel(), single[k][2][0]) and np.array_equal(px.ravel(), single[k][2][1]), k
    compare(single[0], run_oracle(oracle, cs[0]), 1e-8, 1e-7, "coupled patch of the batch")
