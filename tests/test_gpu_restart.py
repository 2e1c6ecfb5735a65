"""Restart of device-resident fields (MPCNSRST v1) through the C ABI.

* the file written from the device holds the device state bit for bit, in the
  layout of the reference's writer (parsed by the independent restatement in
  restart_format.py);
* checkpoint / resume: N steps, restart file, a fresh context reads it, prepares
  and runs M steps; the state equals the uninterrupted N + M steps bit for bit
  (the reference's continuation path, Z4_Mercury/src/MercurySolver.cpp:135-145);
  with the implicit resistive correction the CG's warm start (the previous
  solution) is not part of the file, as in the reference, so the resumed run
  agrees to the CG tolerance (1e-9 relative L2);
* the oracle, loaded from the parsed file, continues identically;
* mismatching files are rejected.
"""
import struct

import numpy as np
import pytest

from mpcns_mercury_amd import cases
from mpcns_mercury_amd.hip import Mercury, MercuryError

import restart_format as rf
from test_gpu_parity import DERIVED, STATE

pytestmark = pytest.mark.gpu

RESTART = ["U_H", "U_Na", "B_xi", "B_eta", "B_zeta"]


def _alloc(case, ib, code):
    d = rf.DELTA_OF_CODE[code]
    return tuple(slice(0, 2 * case.g + m + 1 - dd) for m, dd in zip(case.blocks[ib].dims, d))


def _case(which, integrator=0):
    if which == "shell":
        return cases.make_shell(scale=0.1, integrator=integrator)
    if which == "split":
        return cases.make_split_cube((12, 10, 8), (2, 2, 1), integrator=integrator)
    return cases.make_periodic_cube((10, 8, 6), g=4, resist=0.5, integrator=integrator)


def test_restart_file_holds_the_device_state(tmp_path):
    c = _case("shell")
    g = Mercury(c)
    try:
        g.prepare()
        g.step(3, 0)
        p = tmp_path / "flow_field0000.bin"
        g.write_restart(p, step=3, time=0.125, with_jedge=True)
        step, time, nb, f = rf.read_restart(p)
        assert (step, time, nb) == (3, 0.125, c.nb)
        assert list(f) == RESTART + ["J_xi", "J_eta", "J_zeta"]
        for name, (loc, ncomp, nghost, blocks) in f.items():
            assert nghost == c.g
            for ib, (lo, hi, arr) in enumerate(blocks):
                m = c.blocks[ib].dims
                d = rf.DELTA_OF_CODE[loc]
                assert lo == (-c.g,) * 3 and hi == tuple(m[a] + 1 - d[a] + c.g for a in range(3))
                np.testing.assert_array_equal(arr, g.get(name, ib)[(slice(None),) + _alloc(c, ib, loc)], err_msg=name)
    finally:
        g.close()


@pytest.mark.parametrize("which,integrator,form", [("shell", 0, 1), ("split", 1, 1), ("shell", 0, 2), ("split", 0, 0)])
def test_resume_equals_uninterrupted(tmp_path, which, integrator, form):
    c = _case(which, integrator)
    print(f"[case] {c.name} resume integrator={integrator} form={form}", flush=True)
    a, b = Mercury(c), Mercury(c)
    r = None
    try:
        for x in (a, b):
            x.prepare()
            x.set_form(form)
        dta = a.step(5, integrator)
        dtb = b.step(2, integrator)
        p = tmp_path / "ckpt.bin"
        b.write_restart(p, step=2, time=float(np.sum(dtb)))
        b.close()
        r = Mercury(c)
        step, time = r.read_restart(p)
        assert step == 2 and time == float(np.sum(dtb))
        r.prepare()
        r.set_form(form)
        dtr = r.step(3, integrator)
        np.testing.assert_array_equal(np.concatenate([dtb, dtr]), dta)
        for name in STATE + DERIVED:
            for ib in range(c.nb):
                np.testing.assert_array_equal(r.get(name, ib), a.get(name, ib), err_msg=f"{name}[{ib}]")
    finally:
        a.close()
        if r is not None:
            r.close()


def test_resume_with_resistivity_to_cg_tolerance(tmp_path):
    c = _case("resist")
    a, b = Mercury(c), Mercury(c)
    r = None
    try:
        for x in (a, b):
            x.prepare()
        a.step(4, 0)
        b.step(2, 0)
        p = tmp_path / "ckpt.bin"
        b.write_restart(p, step=2)
        b.close()
        r = Mercury(c)
        r.read_restart(p)
        r.prepare()
        r.step(2, 0)
        for name in STATE:
            num = den = 0.0
            for ib in range(c.nb):
                x, y = r.get(name, ib), a.get(name, ib)
                num += float(np.sum((x - y) ** 2))
                den += float(np.sum(y ** 2))
            assert (num / den) ** 0.5 <= 1e-9, name
    finally:
        a.close()
        if r is not None:
            r.close()


def test_oracle_continues_from_the_device_restart(tmp_path, oracle_mod):
    c = _case("split")
    g = Mercury(c)
    try:
        g.prepare()
        g.step(2, 0)
        p = tmp_path / "ckpt.bin"
        g.write_restart(p)
        _, _, _, f = rf.read_restart(p)
        o = oracle_mod.Oracle(c)
        for name, (loc, ncomp, nghost, blocks) in f.items():
            for ib, (lo, hi, arr) in enumerate(blocks):
                full = o.get(name, ib)
                full[(slice(None),) + _alloc(c, ib, loc)] = arr
                o.set(name, ib, full if ncomp > 1 else full[0])
        o.prepare()
        do = o.step(2, 0)
        dg = g.step(2, 0)
        np.testing.assert_array_equal(dg, do)
        for name in STATE:
            for ib in range(c.nb):
                np.testing.assert_array_equal(g.get(name, ib), o.get(name, ib), err_msg=f"{name}[{ib}]")
    finally:
        g.close()


def test_mismatching_restarts_are_rejected(tmp_path):
    shell, cube = _case("shell"), _case("split")
    g, h = Mercury(shell), Mercury(cube)
    try:
        g.prepare()
        p = tmp_path / "shell.bin"
        g.write_restart(p)
        with pytest.raises(MercuryError):  # block count
            h.read_restart(p)
        # a partial J edge triplet
        _, _, nb, f = rf.read_restart(p)
        q = tmp_path / "partial.bin"
        g.write_restart(q, with_jedge=True)
        _, _, _, fj = rf.read_restart(q)
        loc, ncomp, ng, blocks = fj["J_xi"]
        rf.write_restart(q, 0, 0.0, nb, [("J_xi", loc, ncomp, ng, blocks)])
        with pytest.raises(MercuryError):
            g.read_restart(q)
        # a wrong ghost depth
        loc, ncomp, ng, blocks = f["U_H"]
        rf.write_restart(q, 0, 0.0, nb, [("U_H", loc, ncomp, ng + 1, blocks)])
        with pytest.raises(MercuryError):
            g.read_restart(q)
        # fields this build does not hold are stepped over
        rf.write_restart(q, 7, 1.0, nb, [("T_electron", 0, 1, ng, [(lo, hi, a[:1]) for lo, hi, a in blocks]),
                                         ("U_H", loc, ncomp, ng, blocks)])
        assert g.read_restart(q) == (7, 1.0)
    finally:
        g.close()
        h.close()


def test_truncated_or_oversized_restart_leaves_the_state(tmp_path):
    """A restart whose data stop early, or whose box claims more values than the
    file holds, fails before anything reaches the device: every field keeps the
    values it had, including the fields the file delivered before the fault."""
    shell = _case("shell")
    print(f"[case] {shell.name} truncated restart", flush=True)
    g = Mercury(shell)
    try:
        g.prepare()
        p = tmp_path / "full.bin"
        g.write_restart(p)
        g.step(2, 0)
        before = {n: [g.get(n, ib) for ib in range(shell.nb)] for n in STATE}
        data = p.read_bytes()
        q = tmp_path / "cut.bin"
        q.write_bytes(data[: len(data) - 1000])
        with pytest.raises(MercuryError, match="truncated"):
            g.read_restart(q)
        # the first block box of U_H patched in place: hi[0] grows by 1e6, so the
        # box claims far more values than the file holds
        raw = bytearray(data)
        at = raw.index(b"U_H") + len(b"U_H")
        box_at = at + 12  # location, ncomp, nghost
        hi0 = struct.unpack_from("<i", raw, box_at + 12)[0]
        struct.pack_into("<i", raw, box_at + 12, hi0 + 10**6)
        q.write_bytes(bytes(raw))
        with pytest.raises(MercuryError):
            g.read_restart(q)
        for n in STATE:
            for ib in range(shell.nb):
                np.testing.assert_array_equal(g.get(n, ib), before[n][ib], err_msg=f"{n}[{ib}] changed by a refused restart")
    finally:
        g.close()
