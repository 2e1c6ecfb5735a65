"""Halo plans and alias classes supplied by the caller (MERCURY_BC_EXTERNAL faces).

A context whose block connections are EXTERNAL faces knows no neighbour there;
the caller uploads the halo and owner -> alias copies of every sync group
(mercury_set_sync_plan) and the alias-mean classes of the shared edges
(mercury_set_alias_classes), together with the ghost geometry its own topology
code computes.  Here the plans, classes and geometry are those the library
builds for the same blocks glued by interfaces (read back with
mercury_sync_plan / mercury_alias_classes), so the two contexts must step bit
for bit alike in every form and from a captured graph.  A second test checks
the entry semantics on a hand-built plan with a sign flip and a component
swap, the two things a rotated or mirrored connection needs.
"""
import copy

import numpy as np
import pytest

from mpcns_mercury_amd import cases
from mpcns_mercury_amd.hip import Mercury, MercuryError, field_info

pytestmark = pytest.mark.gpu

GROUPS = ["Ucell", "Bface", "Eedge", "Ehall", "Jedge", "B_cell", "J_cell", "Badd", "Eres", "Eres0", "dB"]
GEOM = ["Jac", "JDxi", "JDet", "JDze", "Area_xi", "Area_eta", "Area_zeta", "beta_xi", "beta_eta", "beta_zeta",
        "dl_xi", "dl_eta", "dl_zeta", "alpha_xi", "alpha_eta", "alpha_zeta", "dr_xi", "dr_eta", "dr_zeta",
        "Bcell_from_Bface_w", "Jcell_from_Jedge_w", "cell_center", "taper_xi", "taper_eta", "taper_zeta",
        "eta_local_xi", "eta_local_eta", "eta_local_zeta"]
INPUTS = ["U_H", "U_Na", "B_xi", "B_eta", "B_zeta", "Badd_xi", "Badd_eta", "Badd_zeta"]
CHECK = ["U_H", "U_Na", "B_xi", "B_eta", "B_zeta", "B_cell", "J_cell", "J_xi", "J_eta", "J_zeta", "E_xi", "PV_H", "U_plus"]


def _external(case):
    c = copy.deepcopy(case)
    for b in c.blocks:
        b.bc = tuple(cases.BC_EXTERNAL if x == cases.BC_INTERFACE else x for x in b.bc)
        b.nbr = (-1,) * 6
        b.off = ((0, 0, 0),) * 6
    return c


def _same(ga, gb, case, ctx):
    for ib in range(case.nb):
        for name in CHECK:
            np.testing.assert_array_equal(gb.get(name, ib), ga.get(name, ib), err_msg=f"{ctx}: {name}[{ib}]")


@pytest.mark.parametrize("form", [0, 1, 2])
def test_caller_plans_reproduce_the_interfaces(form):
    a = cases.make_split_cube(n=(12, 10, 8), split=(2, 2, 1), hall=True)
    b = _external(a)
    print(f"[case] {a.name}: interfaces as EXTERNAL faces with the interface plans, form {form}", flush=True)
    ga, gb = Mercury(a), Mercury(b)
    try:
        for ib in range(a.nb):
            for name in GEOM + INPUTS:
                gb.set(name, ib, ga.get(name, ib))
        n_copies = 0
        for grp in GROUPS:
            for stage in (0, 1):
                plan = ga.sync_plan(grp, stage)
                n_copies += len(plan)
                gb.set_sync_plan(grp, stage, plan)
                np.testing.assert_array_equal(gb.sync_plan(grp, stage), plan)
        assert n_copies > 0
        classes = ga.alias_classes()
        assert len(classes[1]) > 0
        gb.set_alias_classes(*classes)
        for x, y in zip(gb.alias_classes(), classes):
            np.testing.assert_array_equal(x, y)
        for g in (ga, gb):
            g.prepare()
            g.set_form(form)
        np.testing.assert_array_equal(gb.step(3, 0), ga.step(3, 0))
        _same(ga, gb, a, "3 steps")
        if form >= 1:
            for g in (ga, gb):
                g.graph_capture(1)
                g.graph_launch(2)
                g.synchronize()
            _same(ga, gb, a, "graph replay, RK2")
    finally:
        ga.close()
        gb.close()


def test_plan_entry_semantics_and_refusals():
    """dst = sign * src across components and blocks, after the group's own copies."""
    case = _external(cases.make_split_cube(n=(8, 8, 6), split=(2, 1, 1)))
    g = Mercury(case)
    try:
        sx, _, _ = field_info("B_xi")
        sy, _, _ = field_info("B_eta")
        # the first ghost face beyond block 0's -xi face takes minus block 1's
        # eta flux at a chosen owned face, as a mirrored, rotated connection would
        rng = np.random.default_rng(3)
        vals = rng.standard_normal((case.nb,) + case.box(0))
        for ib in range(case.nb):
            g.set("B_eta", ib, vals[ib])
        g.set_sync_plan("Bface", 0, [[sx, 0, -1, 2, 3, sy, 1, 5, 4, 2, -1], [sx, 0, -1, 3, 3, sy, 1, 5, 5, 2, 1]])
        g.op("sync_bface")
        got = g.get("B_xi", 0)[0]
        gg = case.g
        after = g.get("B_eta", 1)[0]
        assert got[gg - 1, gg + 2, gg + 3] == -after[gg + 5, gg + 4, gg + 2]
        assert got[gg - 1, gg + 3, gg + 3] == after[gg + 5, gg + 5, gg + 2]
        with pytest.raises(MercuryError, match="outside the block's box"):
            g.set_sync_plan("Bface", 0, [[sx, 0, -gg - 1, 0, 0, sy, 1, 0, 0, 0, 1]])
        with pytest.raises(MercuryError, match="sign must be"):
            g.set_sync_plan("Bface", 0, [[sx, 0, 0, 0, 0, sy, 1, 0, 0, 0, 2]])
        with pytest.raises(MercuryError, match="unknown sync group"):
            g.set_sync_plan("Bfaces", 0, [])
    finally:
        g.close()
