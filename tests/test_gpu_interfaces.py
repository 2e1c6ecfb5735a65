"""Block interfaces on the GPU, through the C ABI (mercury_create_connected).

The device plan copies each halo ghost straight from the end of the reference's
stage-by-stage hop chain; the oracle runs the stages one hop at a time.  The
two agree bit for bit, and both fill every ghost with the value at its global
position (split_helpers).  Cases: the C1 cube cut into 2 x 2 x 2 blocks (all
faces interfaces, periodicity carried by wrap-around interfaces) and a shell
sector cut into blocks whose interfaces meet the planet surface and the far
field (corner ghosts of mixed inner / physical patches)."""
import numpy as np
import pytest

from mpcns_mercury_amd import cases
from mpcns_mercury_amd.hip import Mercury, MercuryError

from split_helpers import DELTA, N, SPLIT, _box_codes, _classes, _out_axes
from test_gpu_parity import DERIVED, STATE, _cmp, _rel_l2

pytestmark = pytest.mark.gpu


def _run(case, oracle_mod, steps, integrator, form):
    g = Mercury(case)
    o = oracle_mod.Oracle(case)
    g.prepare()
    o.prepare()
    g.set_form(form)
    dg = g.step(steps, integrator)
    do = o.step(steps, integrator)
    return g, o, dg, do


def test_split_sync_fills_global_positions():
    c = cases.make_split_cube(N, SPLIT)
    print(f"[case] {c.name} sync codes", flush=True)
    g = Mercury(c)
    try:
        for ib in range(c.nb):
            for name, dl in (("B_xi", "fx"), ("B_eta", "fy"), ("B_zeta", "fz")):
                g.set(name, ib, _box_codes(c, ib, DELTA[dl]))
            g.set("U_H", ib, np.stack([_box_codes(c, ib, DELTA["cell"], comp=q) for q in range(5)]))
        g.op("sync_bface")
        g.op("sync_ucell")
        for ib in range(c.nb):
            for name, dl in (("B_xi", "fx"), ("B_eta", "fy"), ("B_zeta", "fz")):
                got = g.get(name, ib)[0]
                want = _box_codes(c, ib, DELTA[dl], owned_only=False)
                alloc = tuple(slice(0, 2 * c.g + m + 1 - dd) for m, dd in zip(c.blocks[ib].dims, DELTA[dl]))
                np.testing.assert_array_equal(got[alloc], want[alloc], err_msg=f"{name} block {ib}")
            inb = tuple(slice(0, 2 * c.g + m) for m in c.blocks[ib].dims)
            nout = _out_axes(c, ib, DELTA["cell"])[inb]
            got = g.get("U_H", ib)[(slice(None),) + inb]
            want = np.stack([_box_codes(c, ib, DELTA["cell"], comp=q, owned_only=False) for q in range(5)])[(slice(None),) + inb]
            np.testing.assert_array_equal(got[:, nout <= 1], want[:, nout <= 1])
            assert np.isnan(got[:, nout >= 2]).all()
    finally:
        g.close()


def test_split_owner_alias_and_alias_mean(oracle_mod):
    c = cases.make_split_cube(N, SPLIT)
    print(f"[case] {c.name} owner/alias", flush=True)
    g = Mercury(c)
    o = oracle_mod.Oracle(c)
    try:
        rng = np.random.default_rng(11)
        E = [rng.standard_normal(c.box(ib)) for ib in range(c.nb)]
        tagged = [_box_codes(c, ib, DELTA["fy"], block_tag=True) for ib in range(c.nb)]
        for x in (g, o):
            for ib in range(c.nb):
                x.set("E_zeta", ib, E[ib])
                x.set("B_eta", ib, tagged[ib])
            x.op("alias_e")
            x.op("sync_bface")
        for name in ("E_zeta", "B_eta"):  # NaN outside each location's allocated box, in both
            for ib in range(c.nb):
                np.testing.assert_array_equal(g.get(name, ib), o.get(name, ib), err_msg=f"{name} block {ib}")
        gz = [g.get("E_zeta", ib)[0] for ib in range(c.nb)]
        gb = [g.get("B_eta", ib)[0] for ib in range(c.nb)]
        G = c.g
        for P, mem in _classes(c, DELTA["ez"]).items():
            s = 0.0
            for (b, i, j, k) in mem:
                s += E[b][G + i, G + j, G + k]
            ob, oi, oj, ok = mem[0]
            assert gz[ob][G + oi, G + oj, G + ok] == s / len(mem), P
        for P, mem in _classes(c, DELTA["fy"]).items():
            ob, oi, oj, ok = mem[0]
            for (b, i, j, k) in mem:
                assert gb[b][G + i, G + j, G + k] == tagged[ob][G + oi, G + oj, G + ok]
    finally:
        g.close()


@pytest.mark.parametrize("which", ["cube", "sector_eta", "sector_xi_zeta"])
@pytest.mark.parametrize("integrator", [0, 1])
@pytest.mark.parametrize("form", [0, 1])
def test_split_steps_bit_exact(oracle_mod, which, integrator, form):
    """Steps of split configurations: state and derived fields equal the oracle's bit for bit."""
    if which == "cube":
        c = cases.make_split_cube((16, 12, 10), (2, 2, 2), integrator=integrator)
    elif which == "sector_eta":
        c = cases.make_split_sector(integrator=integrator)
    else:
        c = cases.make_split_sector(cuts=((12, 12), (20,), (8, 8)), integrator=integrator)
    print(f"[case] {c.name} form={form} integrator={integrator}", flush=True)
    g, o, dg, do = _run(c, oracle_mod, 4, integrator, form)
    try:
        np.testing.assert_array_equal(dg, do)
        _cmp(g, o, STATE + DERIVED, f"{c.name} 4 steps")
    finally:
        g.close()


def test_split_hundred_steps_graph(oracle_mod):
    """100 steps of the cut cube replayed from a graph equal the oracle's; div B at roundoff."""
    c = cases.make_split_cube((12, 10, 8), (2, 2, 2), integrator=1)
    print(f"[case] {c.name} 100 steps graph", flush=True)
    g = Mercury(c)
    o = oracle_mod.Oracle(c)
    try:
        g.prepare()
        o.prepare()
        g.set_form(1)
        g.graph_capture(1)
        g.graph_launch(100)
        g.synchronize()
        o.step(100, 1)
        _cmp(g, o, STATE + DERIVED, "100 steps")
        assert g.divb()[0] <= 1e-12
    finally:
        g.close()


def test_split_mixed_accuracy(oracle_mod):
    """Mixed form on the cut shell sector: rel-L2 <= 1e-5 per state variable after 10 steps."""
    c = cases.make_split_sector()
    print(f"[case] {c.name} mixed", flush=True)
    g, o, dg, do = _run(c, oracle_mod, 10, 0, 2)
    try:
        np.testing.assert_allclose(dg, do, rtol=1e-5)
        errs = {n: _rel_l2(g, o, n, c.nb) for n in STATE}
        print("[mixed split] rel-L2:", errs, flush=True)
        assert all(e <= 1e-5 for e in errs.values()), errs
        assert g.divb()[0] <= 1e-12
    finally:
        g.close()


def test_split_resistive_matches_oracle(oracle_mod):
    """Implicit resistive correction with the eta window crossing block interfaces
    (owner / alias unknowns spread over several blocks)."""
    c = cases.make_split_cube((16, 12, 10), (2, 2, 1), resist=0.05)
    print(f"[case] {c.name} resistive", flush=True)
    g, o, dg, do = _run(c, oracle_mod, 3, 0, 1)
    try:
        n, its, rel = g.resist_stats()
        assert n > 0 and its > 0
        for name in STATE:
            assert _rel_l2(g, o, name, c.nb) <= 1e-9, name
    finally:
        g.close()


def test_self_interface_equals_periodic():
    a = Mercury(cases.make_periodic_cube((10, 8, 6)))
    b = Mercury(cases.make_split_cube((10, 8, 6), (1, 1, 1)))
    try:
        for x in (a, b):
            x.prepare()
            x.set_form(1)
            x.step(3, 0)
        for name in STATE + DERIVED:
            np.testing.assert_array_equal(a.get(name, 0), b.get(name, 0), err_msg=name)
    finally:
        a.close()
        b.close()


def test_unsupported_interface_is_rejected():
    c = cases.make_split_cube((10, 9, 8), (2, 1, 1))
    b = c.blocks[0]
    off = list(b.off)
    off[1] = (-3, 0, 0)  # does not glue the +xi face to the neighbour's -xi face
    c.blocks[0] = cases.Block(dims=b.dims, bc=b.bc, nodes=b.nodes, nbr=b.nbr, off=tuple(off))
    with pytest.raises(MercuryError):
        Mercury(c)
