"""The FP32 erfc of the mixed pair kernel (t Q(t) exp(-x^2)) against scipy's
erfc: the fit reproduces from its script, and the kernel's constants in an
FP32 emulation stay within 3e-6 relative over x in [0, 6] (alpha r for every
pair inside the cutoff of the workloads here is below 4.5)."""
import re
import os

import numpy as np
import pytest

from mdbas_amd import erfc_fit

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fit_reproduces_constants():
    c = erfc_fit.fit()
    assert np.allclose(c, erfc_fit.COEFFS, rtol=0, atol=1e-12)


def test_kernel_constants_match_script():
    src = open(os.path.join(HERE, "..", "mdbas_amd", "csrc", "mdg_pair.hip")).read()
    body = src[src.index("mdg_erfc_q(float x"):]
    body = body[:body.index("return")]
    consts = [float(v) for v in re.findall(r"(-?\d+\.\d+)f", body)]
    # fmaf(0.5f, x, 1.0f) first, then the Horner chain from the top coefficient
    assert consts[:2] == [0.5, 1.0]
    horner = consts[2:]
    assert horner[0] == pytest.approx(erfc_fit.COEFFS[-1], abs=0)
    assert horner[1:] == pytest.approx(list(erfc_fit.COEFFS[-2::-1]), abs=0)


def test_fp32_erfc_accuracy():
    from scipy.special import erfc
    x = np.linspace(0.0, 6.0, 300001).astype(np.float32)
    got = erfc_fit.erfc_f32(x).astype(np.float64)
    ref = erfc(x.astype(np.float64))
    assert np.abs(got / ref - 1.0).max() < 3e-6
    # the closed-form points
    assert erfc_fit.erfc_f32(np.float32(0.0)) == pytest.approx(1.0, rel=2e-7)
    assert erfc_fit.q_exact(0.0) == pytest.approx(0.5 / np.sqrt(np.pi), rel=1e-15)


def test_fit_error_is_small_in_fp64():
    tt = np.linspace(0.0, 1.0, 100001)
    q = np.polyval(np.array(erfc_fit.COEFFS)[::-1], tt)
    assert np.abs(q / erfc_fit.q_exact(tt) - 1.0).max() < 2e-7
