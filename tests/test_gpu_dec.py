"""DEC entry points on caller device arrays against the CPU oracle (bit for bit).

Inputs are seeded random fields over the whole block box; every output array
starts as NaN sentinels (and, for Faraday's update, as random faces), so the
comparison also checks that only the owned range of the output is written and
that nothing outside the inputs' owned range (+1) is read.  Shapes cover a
cube, ragged single-cell axes, no ghost layers and a block large enough for
thousands of partial maxima.
"""
import ctypes

import numpy as np
import pytest

import hipmem
import oracle as orc
from mpcns_mercury_amd import hip

pytestmark = pytest.mark.gpu

SHAPES = [((10, 8, 6), 4), ((1, 7, 13), 2), ((9, 1, 1), 0), ((5, 6, 7), 1), ((96, 64, 70), 4)]


def _dims(d):
    return (ctypes.c_int * 3)(*d)


def _chk(rc):
    if rc != 0:
        raise hip.MercuryError(hip.load().mercury_last_error().decode())


def _rand(sh, rng, n):
    return [rng.standard_normal(sh) for _ in range(n)]


@pytest.mark.parametrize("dims,g", SHAPES)
def test_dec_operators_bit_exact(oracle_mod, dims, g):
    print(f"[case] DEC operators dims={dims} g={g}", flush=True)
    L = hip.load()
    sh = orc._box_shape(dims, g)
    rng = np.random.default_rng(sum(dims) + g)
    node = rng.standard_normal(sh)
    edge = _rand(sh, rng, 3)
    face = _rand(sh, rng, 3)
    face0 = _rand(sh, rng, 3)
    nan = np.full(sh, np.nan)
    dn = hipmem.DeviceArray(node)
    de = [hipmem.DeviceArray(x) for x in edge]
    df = [hipmem.DeviceArray(x) for x in face]
    out3 = [hipmem.DeviceArray(nan) for _ in range(3)]
    out1 = hipmem.DeviceArray(nan)
    try:
        d = _dims(dims)
        _chk(L.mercury_dec_d0(d, g, dn.ptr, hipmem.ptrs(out3), None))
        hipmem.synchronize()
        for a, (x, y) in enumerate(zip([o.download() for o in out3], orc.dec_d0(dims, g, node))):
            np.testing.assert_array_equal(x, y, err_msg=f"d0 edge axis {a}")
        for o in out3:
            o.upload(nan)
        _chk(L.mercury_dec_d1(d, g, hipmem.ptrs(de), hipmem.ptrs(out3), None))
        hipmem.synchronize()
        for a, (x, y) in enumerate(zip([o.download() for o in out3], orc.dec_d1(dims, g, edge))):
            np.testing.assert_array_equal(x, y, err_msg=f"d1 face axis {a}")
        _chk(L.mercury_dec_d2(d, g, hipmem.ptrs(df), out1.ptr, None))
        hipmem.synchronize()
        np.testing.assert_array_equal(out1.download(), orc.dec_d2(dims, g, face), err_msg="d2")
        for o, x in zip(out3, face0):
            o.upload(x)
        _chk(L.mercury_dec_faraday_update(d, g, hipmem.ptrs(de), hipmem.ptrs(out3), ctypes.c_double(0.37), None))
        hipmem.synchronize()
        ref = orc.dec_d1(dims, g, edge, out=face0, update=True, dt=0.37)
        for a, (x, y) in enumerate(zip([o.download() for o in out3], ref)):
            np.testing.assert_array_equal(x, y, err_msg=f"faraday face axis {a}")
    finally:
        for x in [dn, out1] + de + df + out3:
            x.free()


@pytest.mark.parametrize("dims,g", SHAPES)
def test_dec_diagnostics_bit_exact(oracle_mod, dims, g):
    """max|d2 d1 E| and max|d1 d0 phi| equal the oracle's sequential maxima
    exactly (a max reduction has no rounding), on random input, on input with
    NaN ghost layers (never read) and on integer input (exactly 0)."""
    print(f"[case] DEC diagnostics dims={dims} g={g}", flush=True)
    L = hip.load()
    sh = orc._box_shape(dims, g)
    rng = np.random.default_rng(3 * sum(dims) + g)
    edge = _rand(sh, rng, 3)
    node = rng.standard_normal(sh) * 1e3
    if g > 0:
        for x in edge:
            x[:g] = np.nan
        node[:, :, :g] = np.nan
    de = [hipmem.DeviceArray(x) for x in edge]
    dn = hipmem.DeviceArray(node)
    res = hipmem.DeviceArray(np.full(1, -1.0))
    try:
        d = _dims(dims)
        _chk(L.mercury_dec_max_abs_div_curl(d, g, hipmem.ptrs(de), res.ptr, None))
        hipmem.synchronize()
        v = res.download()[0]
        assert v == orc.dec_max_abs_div_curl(dims, g, edge) and v > 0.0
        _chk(L.mercury_dec_max_abs_curl_grad(d, g, dn.ptr, res.ptr, None))
        hipmem.synchronize()
        w = res.download()[0]
        assert w == orc.dec_max_abs_curl_grad(dims, g, node) and w > 0.0
        ints = [rng.integers(-99, 99, sh).astype(float) for _ in range(3)]
        for x, y in zip(de, ints):
            x.upload(y)
        _chk(L.mercury_dec_max_abs_div_curl(d, g, hipmem.ptrs(de), res.ptr, None))
        hipmem.synchronize()
        assert res.download()[0] == 0.0
    finally:
        for x in de + [dn, res]:
            x.free()


def test_dec_invalid_arguments_fail_with_a_message():
    """Bad sizes and NULL arrays are refused before anything is launched."""
    L = hip.load()
    one = hipmem.DeviceArray(np.zeros(8))
    try:
        p3 = hipmem.ptrs([one, one, one])
        assert L.mercury_dec_d0(_dims((0, 4, 4)), 2, one.ptr, p3, None) == -1
        assert b"at least one cell" in L.mercury_last_error()
        assert L.mercury_dec_d2(_dims((4, 4, 4)), 2, p3, None, None) == -1
        assert b"NULL" in L.mercury_last_error()
        assert L.mercury_dec_d1(_dims((4, 4, 4)), -1, p3, p3, None) == -1
    finally:
        one.free()
