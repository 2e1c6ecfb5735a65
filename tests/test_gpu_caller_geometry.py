"""Static geometry supplied by the caller.

A caller that computes the static inputs itself (the reference's
METRIC::compute_metric_fields and SetupCellReconstructionWeights_, or any other
code) uploads them over the context's own by catalogue name and calls
mercury_prepare, which refreshes everything derived from them (the FP32 copies
of the mixed form, the derived fields).  Here the uploaded geometry is that of
another periodic mesh with the same block shape, computed on the host by the
library's own geometry routine; the context must then step exactly like one
built on that mesh, in the fused FP64 and the mixed form.
"""
import copy

import numpy as np
import pytest

from mpcns_mercury_amd import cases, hip
from mpcns_mercury_amd.hip import Mercury

pytestmark = pytest.mark.gpu

GEOM = ["Jac", "JDxi", "JDet", "JDze", "Area_xi", "Area_eta", "Area_zeta", "beta_xi", "beta_eta", "beta_zeta",
        "dl_xi", "dl_eta", "dl_zeta", "alpha_xi", "alpha_eta", "alpha_zeta", "dr_xi", "dr_eta", "dr_zeta",
        "Bcell_from_Bface_w", "Jcell_from_Jedge_w", "cell_center", "taper_xi", "taper_eta", "taper_zeta",
        "eta_local_xi", "eta_local_eta", "eta_local_zeta"]
INPUTS = ["U_H", "U_Na", "B_xi", "B_eta", "B_zeta", "Badd_xi", "Badd_eta", "Badd_zeta"]
CHECK = ["U_H", "U_Na", "B_xi", "B_eta", "B_zeta", "B_cell", "J_cell", "PV_H", "U_plus"]


def _rewarped(case):
    """The same block on another smooth periodic mapping (owned nodes moved by
    a displacement of period 1 along every axis)."""
    c = copy.deepcopy(case)
    x = c.blocks[0].nodes
    X, Y, Z = x[0].copy(), x[1].copy(), x[2].copy()
    x[0] = X + 0.02 * np.sin(2 * np.pi * Z)
    x[1] = Y + 0.03 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Z)
    x[2] = Z + 0.015 * np.sin(2 * np.pi * Y)
    return c


@pytest.mark.parametrize("form", [1, 2])
def test_uploaded_geometry_replaces_the_contexts_own(form):
    a = cases.make_periodic_cube(n=(10, 8, 6), g=4, hall=True)
    b = _rewarped(a)
    print(f"[case] {a.name}: geometry of a rewarped mesh uploaded, form {form}", flush=True)
    gb = Mercury(b)
    ga = Mercury(a)
    try:
        moved = False
        for name in GEOM:
            geo = hip.geometry_block(b, 0, name)
            moved = moved or not np.array_equal(geo, ga.get(name, 0))
            ga.set(name, 0, geo)
        assert moved
        # the initial state as the other context holds it before its first
        # synchronisation (B = Bu.JD + d1 A and Badd depend on the mesh; a sync
        # is not idempotent, since aliases it rewrites feed ghosts of the next one)
        for name in INPUTS:
            ga.set(name, 0, gb.get(name, 0))
        ga.prepare()
        gb.prepare()
        for g in (ga, gb):
            g.set_form(form)
        np.testing.assert_array_equal(ga.step(4, 0), gb.step(4, 0))
        for name in CHECK:
            np.testing.assert_array_equal(ga.get(name, 0), gb.get(name, 0), err_msg=name)
    finally:
        ga.close()
        gb.close()
