"""The fast correctly rounded division of the raster kernels on the GPU:
the Newton reciprocal against IEEE 1/b for every |b| in [2^-60, 2^60], and
the Markstein quotient against IEEE a/b on random pairs from the operand
ranges the kernels admit (span quotients, integer row counts, signed zero
numerators).  The frames' bit-exact parity with the oracle rests on it."""
import numpy as np
import pytest

from axi_visualizer_amd import _native

pytestmark = pytest.mark.gpu


def test_fast_division_matches_ieee():
    out = np.zeros(3, np.uint64)
    _native.check(_native.lib().axr_check_division(20260301, 4, _native.ptr(out)),
                  "check_division")
    rcp_bad, tested, div_bad = (int(v) for v in out)
    assert rcp_bad == 0, f"{rcp_bad} reciprocals differ from IEEE 1/b"
    assert tested > 4 * (1 << 26) * 0.8, tested
    assert div_bad == 0, f"{div_bad} of {tested} quotients differ from IEEE a/b"


def test_fast_division_rejects_bad_arguments():
    with pytest.raises(_native.AxrError):
        _native.check(_native.lib().axr_check_division(1, 1, None), "check_division")
    out = np.zeros(3, np.uint64)
    with pytest.raises(_native.AxrError):
        _native.check(_native.lib().axr_check_division(1, 5000, _native.ptr(out)),
                      "check_division")
