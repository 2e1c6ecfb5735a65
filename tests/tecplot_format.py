"""Independent reader of the Tecplot binary files the library writes, for the
tests (not a test module).

Layout of IOModule::WriteTecplotBinFile (src/6_io/IOModule_Tecplot.cpp:691-1017)
in the all-node mode: header (magic, byte order, file type, title, variable
names), one header per ordered zone, the end-of-header marker 357, then per zone
a data section: format flags, passive/shared flags, FP64 min/max per variable
and the FP32 values variable by variable, i fastest.  Strings are one int32 per
character followed by int32 0.
"""
import struct

import numpy as np


class _R:
    def __init__(self, data):
        self.d, self.p = data, 0

    def i32(self):
        v = struct.unpack_from("<i", self.d, self.p)[0]
        self.p += 4
        return v

    def f32(self):
        v = struct.unpack_from("<f", self.d, self.p)[0]
        self.p += 4
        return v

    def f64(self):
        v = struct.unpack_from("<d", self.d, self.p)[0]
        self.p += 8
        return v

    def str(self):
        out = []
        while True:
            c = self.i32()
            if c == 0:
                return "".join(out)
            out.append(chr(c))


def read_plt(path):
    """-> dict(title, names, zones=[dict(name, time, dims=(I, J, K), locs)],
    data=[dict(minmax=[(lo, hi)], values=[array (K, J, I) float32])])"""
    r = _R(open(path, "rb").read())
    assert r.d[:8] == b"#!TDV112"
    r.p = 8
    assert r.i32() == 1  # little endian
    assert r.i32() == 0  # full file
    title = r.str()
    nvar = r.i32()
    names = [r.str() for _ in range(nvar)]
    zones = []
    while True:
        marker = r.f32()
        if marker == 357.0:
            break
        assert marker == 299.0
        name = r.str()
        assert (r.i32(), r.i32()) == (-1, -2)  # parent zone, strand
        time = r.f64()
        assert (r.i32(), r.i32()) == (-1, 0)  # colour, ordered
        assert r.i32() == 1
        locs = [r.i32() for _ in range(nvar)]
        assert (r.i32(), r.i32()) == (0, 0)  # face neighbours
        dims = (r.i32(), r.i32(), r.i32())
        assert r.i32() == 0  # auxiliary data
        zones.append({"name": name, "time": time, "dims": dims, "locs": locs})
    data = []
    for z in zones:
        assert r.f32() == 299.0
        assert [r.i32() for _ in range(nvar)] == [1] * nvar  # float
        assert (r.i32(), r.i32(), r.i32()) == (0, 0, -1)
        minmax = [(r.f64(), r.f64()) for _ in range(nvar)]
        I, J, K = z["dims"]
        n = I * J * K
        values = []
        for _ in range(nvar):
            a = np.frombuffer(r.d, dtype="<f4", count=n, offset=r.p).reshape(K, J, I)
            r.p += 4 * n
            values.append(a)
        data.append({"minmax": minmax, "values": values})
    assert r.p == len(r.d)
    return {"title": title, "names": names, "zones": zones, "data": data}


def node_average(field, jac, g, dims):
    """EvalValue_CellToNode_ (src/6_io/IOModule_Tecplot.cpp:322-418) restated in
    numpy over the owned nodes of one block: |Jac|-weighted mean of the 8 cells
    around each node, in the reference's order; returns (m0+1, m1+1, m2+1)."""
    m0, m1, m2 = dims
    wsum = np.zeros((m0 + 1, m1 + 1, m2 + 1))
    w = np.zeros_like(wsum)
    for dk in (-1, 0):
        for dj in (-1, 0):
            for di in (-1, 0):
                sl = tuple(slice(g + d, g + d + m + 1) for d, m in zip((di, dj, dk), dims))
                wt = np.abs(jac[sl])
                ok = np.isfinite(wt) & (wt > 0.0)
                wsum = np.where(ok, wsum + wt * field[sl], wsum)
                w = np.where(ok, w + wt, w)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(w > 0.0, wsum / np.where(w > 0.0, w, 1.0), 0.0)
