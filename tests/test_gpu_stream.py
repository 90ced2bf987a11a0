"""The device's HBM streaming rates (gd_hbm_stream_rates): a copy and a read of
a buffer through the library's own 16 B-vector kernels, the figures bench.py
sets the step's bandwidth beside.  The rates must be positive and below what
the device's HBM can physically deliver (the published 8 TB/s, with margin for
a box that runs above it), for a buffer well past the caches and for one of
the smallest accepted size."""
import pytest

from conftest import pkg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes", [1 << 30, 16 * 1024])
def test_stream_rates_are_plausible(nbytes):
    L = pkg()._lib
    copy, read = L.hbm_stream_rates(nbytes, 5)
    assert copy > 0.0 and read > 0.0
    if nbytes >= 1 << 30:
        # out of every cache: bounded by HBM, and well above a broken kernel
        assert 1000.0 < copy < 10000.0 and 1000.0 < read < 10000.0
