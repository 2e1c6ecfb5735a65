"""The no-allocation form of max|d2 d1 E| against the allocating one (bit for bit).

mercury_dec_max_abs_div_curl_noalloc keeps its partial maxima in the library's
static device array instead of a stream-ordered allocation.  Two calls in a row,
the second on a smaller block (the array keeps the first call's partials, which
the second must not see), each equal to the allocating entry point and to the
oracle; a block beyond the array's limit is refused before any launch.
"""
import ctypes

import numpy as np
import pytest

import hipmem
import oracle as orc
from mpcns_mercury_amd import hip

pytestmark = pytest.mark.gpu


def _dims(d):
    return (ctypes.c_int * 3)(*d)


def _chk(rc):
    if rc != 0:
        raise hip.MercuryError(hip.load().mercury_last_error().decode())


def _one_call(L, dims, g, seed, scale):
    sh = orc._box_shape(dims, g)
    rng = np.random.default_rng(seed)
    edge = [rng.standard_normal(sh) * scale for _ in range(3)]
    de = [hipmem.DeviceArray(x) for x in edge]
    r0 = hipmem.DeviceArray(np.full(1, -1.0))
    r1 = hipmem.DeviceArray(np.full(1, -2.0))
    try:
        _chk(L.mercury_dec_max_abs_div_curl_noalloc(_dims(dims), g, hipmem.ptrs(de), r0.ptr, None))
        _chk(L.mercury_dec_max_abs_div_curl(_dims(dims), g, hipmem.ptrs(de), r1.ptr, None))
        hipmem.synchronize()
        a, b = r0.download()[0], r1.download()[0]
    finally:
        for x in de + [r0, r1]:
            x.free()
    return a, b, orc.dec_max_abs_div_curl(dims, g, edge)


def test_noalloc_equals_allocating_entry_point(oracle_mod):
    L = hip.load()
    print("[case] noalloc div-curl: 128x96x80 then 6x5x4", flush=True)
    # first call: large values on a large block fill many partial slots
    a, b, o = _one_call(L, (128, 96, 80), 4, 11, 1e6)
    assert a == b == o and a > 0.0
    # second call: a smaller block with values 1e6 times smaller; a stale partial
    # of the first call would dominate the maximum
    a2, b2, o2 = _one_call(L, (6, 5, 4), 2, 12, 1.0)
    assert a2 == b2 == o2 and 0.0 < a2 < a * 1e-3


def test_noalloc_refuses_blocks_beyond_its_scratch():
    L = hip.load()
    one = hipmem.DeviceArray(np.zeros(8))
    try:
        # 512^3 cells = 524 288 partials > 131 072: refused on the host, nothing launched
        rc = L.mercury_dec_max_abs_div_curl_noalloc(_dims((512, 512, 512)), 1, hipmem.ptrs([one, one, one]), one.ptr, None)
        assert rc == -1
        assert b"scratch limit" in L.mercury_last_error()
    finally:
        one.free()
