"""The device arithmetic of the raster kernels beyond plain IEEE operations,
on the GPU: the Newton reciprocal against IEEE 1/b for every |b| in
[2^-60, 2^60]; the Markstein quotient against IEEE a/b on random pairs from
the operand ranges the kernels admit (span quotients, integer row counts,
signed zero numerators); the hardware's saturating float-to-int conversions
against their formulas for every float.  The frames' bit-exact parity with
the oracle rests on them."""
import numpy as np
import pytest

from axi_visualizer_amd import _native

pytestmark = pytest.mark.gpu


def test_device_arithmetic_matches_ieee_and_formulas():
    out = np.zeros(4, np.uint64)
    _native.check(_native.lib().axr_check_arith(20260301, 4, _native.ptr(out)), "check_arith")
    rcp_bad, tested, div_bad, cvt_bad = (int(v) for v in out)
    assert rcp_bad == 0, f"{rcp_bad} reciprocals differ from IEEE 1/b"
    assert tested > 4 * (1 << 26) * 0.8, tested
    assert div_bad == 0, f"{div_bad} of {tested} quotients differ from IEEE a/b"
    assert cvt_bad == 0, f"{cvt_bad} saturating conversions differ from their formulas"


def test_arith_check_rejects_bad_arguments():
    with pytest.raises(_native.AxrError):
        _native.check(_native.lib().axr_check_arith(1, 1, None), "check_arith")
    out = np.zeros(4, np.uint64)
    with pytest.raises(_native.AxrError):
        _native.check(_native.lib().axr_check_arith(1, 5000, _native.ptr(out)), "check_arith")
