"""jsv_max_abs_diff_noalloc: the distance (max |a - b|) without any device
allocation, its result words in a fixed device array.  Checked through the
C ABI against jsv_max_abs_diff (bit for bit: a max is exact), on calls in a
row where a later call is smaller than an earlier one (the array keeps the
last call's contents), per segment, and at the segment limit."""
import ctypes

import numpy as np
import pytest
import torch

from jsolver_amd import _lib
from jsolver_amd._device import ptr, stream_ptr

pytestmark = pytest.mark.gpu


def _ref(a, b, n):
    out = ctypes.c_double()
    _lib.check(_lib.load().jsv_max_abs_diff(ptr(a), ptr(b), n, ctypes.byref(out), stream_ptr()))
    return out.value


def _noalloc(a, b, n, nseg):
    out = np.zeros(nseg)
    _lib.check(_lib.load().jsv_max_abs_diff_noalloc(
        ptr(a), ptr(b), n, nseg, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), stream_ptr()))
    return out


def _pair(n, seed, scale=1.0):
    g = torch.Generator(device="cpu").manual_seed(seed)
    a = (scale * torch.randn(n, generator=g, dtype=torch.float64)).cuda()
    b = (scale * torch.randn(n, generator=g, dtype=torch.float64)).cuda()
    return a, b


def test_two_calls_second_smaller():
    # a large difference first, then a smaller input whose maximum is lower:
    # a word left over from the first call would show
    a1, b1 = _pair(1 << 22, 1, scale=100.0)
    a2, b2 = _pair(1000, 2, scale=0.01)
    for a, b in ((a1, b1), (a2, b2)):
        n = a.numel()
        got = _noalloc(a, b, n, 1)[0]
        want = _ref(a, b, n)
        assert got == want
        assert got == float(torch.max(torch.abs(a - b)))


def test_segments_match_per_segment_calls():
    nseg, n = 37, 5003
    a, b = _pair(nseg * n, 3)
    got = _noalloc(a, b, n, nseg)
    want = np.array([_ref(a[s * n:], b[s * n:], n) for s in range(nseg)])
    np.testing.assert_array_equal(got, want)
    # then fewer segments of a smaller size
    got2 = _noalloc(a, b, 100, 5)
    want2 = np.array([_ref(a[s * 100:], b[s * 100:], 100) for s in range(5)])
    np.testing.assert_array_equal(got2, want2)


def test_segment_limit_and_nan():
    nmax = _lib.JSV_NOALLOC_MAX_SEGMENTS
    a, b = _pair(nmax * 64, 4)
    got = _noalloc(a, b, 64, nmax)
    want = torch.max(torch.abs(a - b).view(nmax, 64), dim=1).values.cpu().numpy()
    np.testing.assert_array_equal(got, want)
    with pytest.raises(_lib.JsvError, match="at most"):
        _noalloc(a, b, 64, nmax + 1)
    # NaN anywhere in a segment gives NaN there, as jsv_max_abs_diff
    a[70] = float("nan")
    got = _noalloc(a, b, 64, 3)
    assert np.isnan(got[1]) and not np.isnan(got[0]) and not np.isnan(got[2])
    assert np.isnan(_ref(a[64:], b[64:], 64))
    # empty segments give 0
    assert _noalloc(a, b, 0, 4).tolist() == [0.0] * 4
