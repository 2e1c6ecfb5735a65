"""Independent restatement of the reference's restart and run-data files, for
the tests (not a test module).

MPCNSRST version 1: IOModule::WriteRestartBinFile / ReadRestartBinFile,
src/6_io/IOModule_Bin.cpp:168-420 (strings: i32 length + bytes,
src/6_io/IOModule_Bin.cpp:39-60; location codes :62-84).
MPCNSRUN version 1: RunData::WriteBinary / ReadBinary, include/6_io/RunData.h:98-205.
"""
import struct

import numpy as np

LOC_CODE = {"cell": 0, "fx": 1, "fy": 2, "fz": 3, "ex": 4, "ey": 5, "ez": 6, "node": 7}
DELTA_OF_CODE = {0: (1, 1, 1), 1: (0, 1, 1), 2: (1, 0, 1), 3: (1, 1, 0), 4: (1, 0, 0), 5: (0, 1, 0), 6: (0, 0, 1), 7: (0, 0, 0)}


def _str(s):
    b = s.encode()
    return struct.pack("<i", len(b)) + b


class _Reader:
    def __init__(self, data):
        self.d, self.p = data, 0

    def take(self, fmt):
        v = struct.unpack_from(fmt, self.d, self.p)
        self.p += struct.calcsize(fmt)
        return v if len(v) > 1 else v[0]

    def str(self):
        n = self.take("<i")
        s = self.d[self.p:self.p + n].decode()
        self.p += n
        return s

    def f64(self, n):
        a = np.frombuffer(self.d, dtype="<f8", count=n, offset=self.p).copy()
        self.p += 8 * n
        return a


def read_restart(path):
    """-> (step, time, nblock, {name: (loc, ncomp, nghost, [per block (lo, hi, array (ncomp, ni, nj, nk) or None)])})"""
    r = _Reader(open(path, "rb").read())
    assert r.d[:8] == b"MPCNSRST"
    r.p = 8
    version, step, time, nblock, nfield = r.take("<i"), r.take("<i"), r.take("<d"), r.take("<i"), r.take("<i")
    assert version == 1
    fields = {}
    for _ in range(nfield):
        name = r.str()
        loc, ncomp, nghost = r.take("<i"), r.take("<i"), r.take("<i")
        blocks = []
        for _ in range(nblock):
            lo = r.take("<3i")
            hi = r.take("<3i")
            active = r.take("<i")
            if not active:
                blocks.append((lo, hi, None))
                continue
            n = [hi[a] - lo[a] for a in range(3)]
            vals = r.f64(n[0] * n[1] * n[2] * ncomp).reshape(n[0], n[1], n[2], ncomp)
            blocks.append((lo, hi, np.moveaxis(vals, 3, 0)))
        fields[name] = (loc, ncomp, nghost, blocks)
    assert r.p == len(r.d)
    return step, time, nblock, fields


def write_restart(path, step, time, nblock, fields):
    """fields: list of (name, loc, ncomp, nghost, [per block (lo, hi, array (ncomp, ni, nj, nk) or None)])"""
    out = [b"MPCNSRST", struct.pack("<iidii", 1, step, time, nblock, len(fields))]
    for name, loc, ncomp, nghost, blocks in fields:
        out.append(_str(name) + struct.pack("<iii", loc, ncomp, nghost))
        for lo, hi, arr in blocks:
            out.append(struct.pack("<3i", *lo) + struct.pack("<3i", *hi) + struct.pack("<i", 0 if arr is None else 1))
            if arr is not None:
                out.append(np.ascontiguousarray(np.moveaxis(arr, 0, 3), dtype="<f8").tobytes())
    open(path, "wb").write(b"".join(out))


def read_rundata(path):
    r = _Reader(open(path, "rb").read())
    assert r.d[:8] == b"MPCNSRUN"
    r.p = 8
    assert r.take("<i") == 1
    step, time, dt = r.take("<i"), r.take("<d"), r.take("<d")
    nref = r.take("<i")
    ref = r.f64(nref)
    rmax = r.take("<d")
    md = {}
    for _ in range(r.take("<i")):
        k = r.str()
        md[k] = r.take("<d")
    mi = {}
    for _ in range(r.take("<i")):
        k = r.str()
        mi[k] = r.take("<i")
    assert r.p == len(r.d)
    return step, time, dt, ref, rmax, md, mi


def write_rundata(path, step, time, dt, ref, rmax, md=None, mi=None):
    md, mi = md or {}, mi or {}
    out = [b"MPCNSRUN", struct.pack("<iidd", 1, step, time, dt), struct.pack("<i", len(ref)),
           np.asarray(ref, dtype="<f8").tobytes(), struct.pack("<d", rmax), struct.pack("<i", len(md))]
    for k in sorted(md):
        out.append(_str(k) + struct.pack("<d", md[k]))
    out.append(struct.pack("<i", len(mi)))
    for k in sorted(mi):
        out.append(_str(k) + struct.pack("<i", mi[k]))
    open(path, "wb").write(b"".join(out))
