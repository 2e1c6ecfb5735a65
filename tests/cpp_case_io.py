"""Input and output files of tests/cpp/mercury_caller (the C++ drop-in caller)."""
import ctypes
import struct

import numpy as np

STATE = (("U_H", 5), ("U_Na", 5), ("B_xi", 1), ("B_eta", 1), ("B_zeta", 1))


def write_input(path, case, fields):
    """fields: list over blocks of {name: (ncomp, box...) array} uploaded after the solver's setup."""
    nb, g = case.nb, case.g
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", nb, g))
        f.write(case.dims_array().astype("<i4").tobytes())
        f.write(case.bc_array().astype("<i4").tobytes())
        f.write(case.nbr_array().astype("<i4").tobytes())
        f.write(case.trans_array().astype("<i4").tobytes())
        pt = np.zeros((0, 18), np.int32) if case.patches is None else np.asarray(case.patches, dtype=np.int32).reshape(-1, 18)
        f.write(struct.pack("<ii", len(pt), int(case.patch_rule)))
        f.write(pt.astype("<i4").tobytes())
        f.write(case.nodes_flat().astype("<f8").tobytes())
        f.write(bytes(ctypes.string_at(ctypes.addressof(case.phys), ctypes.sizeof(case.phys))))
        dip = case.dipole_array()
        f.write(struct.pack("<i", len(dip)))
        f.write(dip.astype("<f8").tobytes())
        names = list(fields[0].keys())
        f.write(struct.pack("<i", len(names)))
        for name in names:
            nc = fields[0][name].shape[0]
            f.write(struct.pack("<i", len(name)))
            f.write(name.encode())
            f.write(struct.pack("<i", nc))
            for ib in range(nb):
                arr = np.ascontiguousarray(fields[ib][name], dtype="<f8")
                assert arr.shape == (nc,) + case.box(ib), (name, arr.shape)
                f.write(arr.tobytes())


def read_output(path, case, nsteps):
    raw = np.fromfile(path, dtype="<f8")
    dts, rest = raw[:nsteps], raw[nsteps:]
    divb, div_curl, curl_grad = rest[:3]
    pos = 3
    state = {}
    for name, nc in STATE:
        per_block = []
        for ib in range(case.nb):
            n = nc * int(np.prod(case.box(ib)))
            per_block.append(rest[pos:pos + n].reshape((nc,) + case.box(ib)))
            pos += n
        state[name] = per_block
    assert pos == len(rest)
    return dts, divb, div_curl, curl_grad, state
