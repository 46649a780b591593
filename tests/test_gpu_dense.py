"""The dense FP64-MFMA influence product (k_dense) of patches of at most 2500 elements, through
the cntc_* C ABI.  It is the reference's direct summation (src/m_aijpj.f90:99-254, 424-448)
and equals the FFT product to rounding.

CNTC_DENSE selects where it runs: 1 (default) in the sensitivity solves, 2 in every product of
a small patch, 0 nowhere.  With every product dense the solves match the oracle to the same
bar as the FFT path: at eps 1e-12, Fn, Fx, Fy within 1e-9 of themselves (Fy of mu Fn where it
vanishes), tractions within 1e-8 of each component's largest value, identical element states.
Dense and FFT solves of the same case agree to the same bar; sensitivities with tight
sensitivity solves (epsens 1e-10) agree within 1e-7 of each column's largest entry.
"""
import numpy as np
import pytest

from tests import cases as C
from tests.test_gpu_parity import run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def force_err(f_a, f_b, mu=0.3):
    f_a, f_b = np.array(f_a[:3]), np.array(f_b[:3])
    den = np.maximum(np.abs(f_b), mu * abs(f_b[0]) * np.array([0.0, 1e-3, 1e-3]))
    den[0] = abs(f_b[0])
    return np.abs(f_a - f_b) / den


def traction_err(p_a, p_b):
    return max(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300) for a, b in zip(p_a, p_b))


CASES = {
    # T=1 shift by TangCG (products of cs and of the preconditioner), 17 x 17
    "shift 17x17 TangCG": dict(C.cattaneo_case(), eps=1e-12, maxgs=2000),
    # T=3 steady rolling by GDsteady, ragged tiles (37 x 29)
    "rolling 37x29 GDsteady": C.hertz_case(6.0, 4.5, 37, 29, 8.0e4, cksi=-0.0008, ceta=0.0002,
                                           cphi=0.0005, eps=1e-12, maxgs=2000, maxin=30),
    # T=2 transient rolling from rest by TangCG, 50 x 50 = 2500 elements (the largest size)
    "transient 50x50 TangCG": C.hertz_case(6.0, 6.0, 50, 50, 8.2e3, cksi=0.0003, eps=1e-12,
                                           maxgs=2000, maxin=30, tang=2),
}


@pytest.mark.parametrize("name", list(CASES))
def test_dense_products_vs_oracle_and_fft(cntc, oracle, monkeypatch, name):
    c = CASES[name]
    monkeypatch.setenv("CNTC_DENSE", "2")
    g = run_gpu(cntc, c, ire=611)
    monkeypatch.setenv("CNTC_DENSE", "0")
    f = run_gpu(cntc, c, ire=612)
    monkeypatch.delenv("CNTC_DENSE")
    o = run_oracle(oracle, c)
    assert g[0] == f[0] == o[0] == 0
    ef, ep = force_err(g[1], o[1]), traction_err(g[2], o[2])
    ef2, ep2 = force_err(g[1], f[1]), traction_err(g[2], f[2])
    print(name, "dense", g[1], "fft", f[1], "oracle", o[1], "\n  vs oracle: force", ef,
          "traction %.2e" % ep, "| vs fft: force", ef2, "traction %.2e" % ep2, flush=True)
    assert ef.max() <= 1e-9 and ep <= 1e-8
    assert ef2.max() <= 1e-9 and ep2 <= 1e-8
    assert np.array_equal(g[3], o[3]) and np.array_equal(g[3], f[3])


def test_dense_sensitivities_equal_fft(cntc, oracle, monkeypatch):
    """S=3 on a 30 x 30 steady-rolling patch with tight sensitivity solves: the dense path
    (default) and the FFT path give the same 4 x 4 sensitivities, and both match the oracle."""
    c = C.hertz_case(6.0, 4.0, 30, 30, 8200.0, cksi=0.0004, ceta=0.0002, cphi=0.001, eps=1e-10,
                     maxgs=2000, maxin=30)
    c.update(sens=3, mxsens=200, epsens=1e-10)
    out = {}
    for mode, ire in (("1", 613), ("0", 614)):
        monkeypatch.setenv("CNTC_DENSE", mode)
        C.apply_cntc(cntc, ire, 1, c)
        assert cntc.calculate(ire, 1) == 0
        out[mode] = cntc.getsensitivities(ire, 1)
    monkeypatch.delenv("CNTC_DENSE")
    P = oracle.Problem()
    assert P.calculate(*C.to_oracle(oracle, c)) == 0
    S_o = P.sens()
    print("dense\n", out["1"], "\nfft\n", out["0"], "\noracle\n", S_o, flush=True)
    for k in range(4):
        scale = max(np.abs(S_o[:, k]).max(), 1e-300)
        assert np.abs(out["1"][:, k] - out["0"][:, k]).max() <= 1e-7 * scale, k
        assert np.abs(out["1"][:, k] - S_o[:, k]).max() <= 1e-7 * scale, k


def test_dense_kernel_runs_in_sensitivity_solves(cntc):
    """The default routes the sensitivity solves of a small patch through k_dense (its launches
    are counted by the per-kernel timing) and leaves the main solve on the FFT products."""
    c = C.hertz_case(6.0, 4.0, 24, 24, 8200.0, cksi=0.0004, cphi=0.001, eps=1e-8, maxgs=999,
                     maxin=30)
    C.apply_cntc(cntc, 615, 1, c)
    cntc.amdprofile(True)
    assert cntc.calculate(615, 1) == 0
    plain = cntc.amdprofileget()
    c["sens"] = 3
    C.apply_cntc(cntc, 616, 1, c)
    cntc.amdprofile(True)
    assert cntc.calculate(616, 1) == 0
    sens = cntc.amdprofileget()
    cntc.amdprofile(False)
    assert "dense" not in plain
    assert sens["dense"][2] > 0
