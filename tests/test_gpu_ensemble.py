"""Ensemble of independent runs stepped by one graph (the batched configuration
at small sizes): every member equals its own oracle bit for bit."""
import numpy as np
import pytest

from mpcns_mercury_amd import cases
from mpcns_mercury_amd.hip import Ensemble, Mercury

pytestmark = pytest.mark.gpu

STATE = ["U_H", "U_Na", "B_xi", "B_eta", "B_zeta"]


@pytest.mark.parametrize("integrator", [0, 1])
def test_ensemble_members_bit_exact(oracle_mod, integrator):
    runs = [cases.make_ensemble_run(r, scale=0.12) for r in range(3)]
    for c in runs:
        print(f"[case] {c.name} dims={[b.dims for b in c.blocks]} hall={c.phys.hall_coef:.3f}", flush=True)
    members = [Mercury(c) for c in runs]
    oracles = [oracle_mod.Oracle(c) for c in runs]
    try:
        for m, o in zip(members, oracles):
            m.prepare()
            m.set_form(1)
            o.prepare()
        ens = Ensemble(members, integrator)
        ens.launch(3)
        ens.close()
        for c, m, o in zip(runs, members, oracles):
            o.step(3, integrator)
            for name in STATE:
                for ib in range(c.nb):
                    np.testing.assert_array_equal(m.get(name, ib), o.get(name, ib), err_msg=f"{c.name} {name}[{ib}]")
            assert m.divb()[0] <= 1e-12
    finally:
        for m in members:
            m.close()


def test_ensemble_runs_differ():
    """The seeded members differ in block count or sizes, inflow, IMF and Hall coefficient."""
    runs = [cases.make_ensemble_run(r) for r in range(8)]
    assert len({tuple(b.dims for b in c.blocks) for c in runs}) == 8
    assert len({round(c.phys.hall_coef, 6) for c in runs}) == 8
    assert all(0.02 <= c.phys.hall_coef <= 0.1 for c in runs)
    assert all(1 <= c.nb <= 4 and all(32 <= d <= 96 for b in c.blocks for d in b.dims) for c in runs)
