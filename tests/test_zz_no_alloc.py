"""GPU test of fehip_group_device_noalloc (no reference counterpart): the
grouping without allocation, against fehip_group (src/FixedEffect.jl:18-21's
GroupedArray numbering as this library defines it) on the same keys, through
the C-ABI -- two calls in a row, the second with a smaller input (the shared
scratch keeps the first call's contents), plus the size limit."""
import ctypes as C
import sys

import numpy as np
import pytest

from fixedeffects_rocm import api, device
from fixedeffects_rocm._lib import FEHIP_EDIM, lib

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(cols):
    N = len(cols[0])
    bufs = []
    for c in cols:
        b = device.DeviceBuffer(N)
        b.upload(np.ascontiguousarray(c, np.int64).view(np.float64))
        bufs.append(b)
    out = device.DeviceBuffer(N)
    ng = C.c_int64(-1)
    arr = (C.c_void_p * len(cols))(*[C.c_void_p(b.address) for b in bufs])
    rc = lib.fehip_group_device_noalloc(N, len(cols), arr, C.c_void_p(out.address), C.byref(ng))
    refs = out.download(N).view(np.int64).copy() if rc == 0 else None
    return rc, refs, int(ng.value)


def test_noalloc_grouping_matches_fehip_group(gpu):
    rng = np.random.default_rng(8)
    cases = [
        # N = 2^20 (the maximum), two columns, keys spanning the int64 range
        [rng.integers(-2**62, 2**62, 1 << 20), rng.integers(0, 50, 1 << 20)],
        # then a smaller call: one column of few distinct keys
        [rng.integers(-5, 7, 70_001)],
    ]
    for cols in cases:
        N = len(cols[0])
        _say(f"case: noalloc grouping N={N} ncols={len(cols)}")
        want, nwant = api.group(*cols)
        rc, got, ng = _noalloc(cols)
        assert rc == 0, lib.fehip_last_error()
        assert ng == nwant
        np.testing.assert_array_equal(got, want)
    # beyond the maximum: a clear error, nothing run
    _say("case: noalloc grouping N=2^20+1")
    rc, got, ng = _noalloc([np.zeros((1 << 20) + 1, np.int64)])
    assert rc == FEHIP_EDIM and b"maximum" in lib.fehip_last_error()
