"""Independent runs in ONE context (mercury_create_ensemble): every pass of a
step is one launch over every block of every run, each run with its own
physics (Hall coefficient, inflow, IMF), CFL time step and abort status.  Each
run's state and dt equal its own CPU oracle run bit for bit."""
import numpy as np
import pytest

from mpcns_mercury_amd import cases
from mpcns_mercury_amd.hip import Mercury, MercuryError

pytestmark = pytest.mark.gpu

STATE = ["U_H", "U_Na", "B_xi", "B_eta", "B_zeta"]


def _runs(n=3, scale=0.25):
    return [cases.make_ensemble_run(r, scale=scale) for r in range(n)]


def _check_runs(g, runs, oracles, ctx, exact=True):
    base = 0
    for r, (c, o) in enumerate(zip(runs, oracles)):
        for ib in range(c.nb):
            for name in STATE:
                a, b = g.get(name, base + ib), o.get(name, ib)
                if exact:
                    np.testing.assert_array_equal(a, b, err_msg=f"{ctx}: run {r} block {ib} {name}")
                else:
                    assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b), f"{ctx}: run {r} block {ib} {name}"
        base += c.nb


@pytest.mark.parametrize("integrator", [0, 1])
@pytest.mark.parametrize("form", [0, 1])
def test_ensemble_context_equals_each_runs_oracle(oracle_mod, form, integrator):
    runs = _runs()
    print(f"[case] ensemble context {len(runs)} runs {[c.nb for c in runs]} form={form} integrator={integrator}", flush=True)
    g = Mercury(cases.concat_runs(runs))
    oracles = [oracle_mod.Oracle(c) for c in runs]
    try:
        g.prepare()
        g.set_form(form)
        for o in oracles:
            o.prepare()
        odts = [o.step(4, integrator) for o in oracles]
        g.step(4, integrator, record_dt=False)
        np.testing.assert_array_equal(g.run_dts(), [d[-1] for d in odts])
        assert len(set(g.run_dts())) == len(runs)  # each run keeps its own dt
        _check_runs(g, runs, oracles, f"form {form}")
        assert g.divb()[0] <= 1e-12
    finally:
        g.close()


def test_ensemble_context_graph_and_mixed(oracle_mod):
    runs = _runs(4)
    print(f"[case] ensemble context {len(runs)} runs graph + mixed", flush=True)
    oracles = [oracle_mod.Oracle(c) for c in runs]
    for o in oracles:
        o.prepare()
        o.step(10, 0)
    for form in (1, 2):
        g = Mercury(cases.concat_runs(runs))
        try:
            g.prepare()
            g.set_form(form)
            g.graph_capture(0)
            g.graph_launch(10)
            g.synchronize()
            _check_runs(g, runs, oracles, f"graph form {form}", exact=(form == 1))
        finally:
            g.close()


def test_ensemble_context_abort_names_the_run(oracle_mod):
    """A runaway cell in run 1 stops run 1 only: the call fails naming run 1,
    run 1 keeps its last completed state and the other runs advance."""
    runs = _runs()
    print("[case] ensemble context dt abort in run 1", flush=True)
    g = Mercury(cases.concat_runs(runs))
    try:
        b1 = runs[0].nb  # first block of run 1
        U = g.get("U_H", b1)
        i = runs[1].g + 2
        U[1, i, i, i] = 1e10 * U[0, i, i, i]
        g.set("U_H", b1, U)
        g.prepare()
        g.op("sync_ucell")  # the planted value's ghosts as the step's own fill leaves them
        g.set_form(1)
        before = [g.get("U_H", b) for b in range(g.case.nb)]
        with pytest.raises(MercuryError, match="run 1: dt too small"):
            g.step(2, 0, record_dt=False)
        for b in range(runs[0].nb, runs[0].nb + runs[1].nb):
            np.testing.assert_array_equal(g.get("U_H", b), before[b])
        assert not np.array_equal(g.get("U_H", 0), before[0])
        dts = g.run_dts()
        assert dts[1] == 0.0 and dts[0] > 0.0 and dts[2] > 0.0
    finally:
        g.close()
