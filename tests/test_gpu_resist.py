"""Implicit resistive correction on the GPU against the CPU oracle (through the C ABI).

The operator is evaluated with the oracle's operations in the oracle's order,
so one application is compared exactly.  The CG solutions of the two sides
stop at the relative residual 1e-10 from different rounding paths; they are
compared at 1e-8 relative to the largest value, and the state after steps with
the correction at 1e-9 relative L2.
"""
import numpy as np
import pytest

from mpcns_mercury_amd import cases
from mpcns_mercury_amd.hip import Mercury, MercuryError

pytestmark = pytest.mark.gpu

AX = ("_xi", "_eta", "_zeta")


def _case(shape):
    if shape == "cube":
        return cases.make_periodic_cube(n=(10, 8, 6), g=4, resist=0.5)
    if shape == "rshell":  # 2 x 2 blocks glued by interfaces, exponential eta on the whole shell
        return cases.make_resistive_shell(scale=0.06)
    return cases.make_shell(scale=0.1, resist=0.5)


@pytest.fixture(params=["cube", "shell", "rshell"])
def rpair(request, oracle_mod):
    case = _case(request.param)
    print(f"[case] {case.name} dims={[b.dims for b in case.blocks]}", flush=True)
    g = Mercury(case)
    o = oracle_mod.Oracle(case)
    g.prepare()
    o.prepare()
    yield case, g, o
    g.close()


def test_unknowns_and_eta_bit_exact(rpair):
    case, g, o = rpair
    n, _, _ = g.resist_stats()
    assert n > 50
    for ib in range(case.nb):
        for s in AX:
            np.testing.assert_array_equal(g.get("eta_res" + s, ib), o.get("eta_res" + s, ib))


@pytest.mark.parametrize("op", ["res_matvec", "res_matvec_lin"])
def test_operator_bit_exact(rpair, op):
    case, g, o = rpair
    dt = o.timestep()
    rng = np.random.default_rng(11)
    for ib in range(case.nb):
        for s in AX:
            x = rng.standard_normal(case.box(ib))[None]
            g.set("edge" + s, ib, x)
            o.set("edge" + s, ib, x)
    g.op(op, dt)
    o.op(op, dt)
    for ib in range(case.nb):
        for s in AX:
            np.testing.assert_array_equal(g.get("edge" + s, ib), o.get("edge" + s, ib))


def test_solve_matches_oracle(rpair):
    case, g, o = rpair
    dt = o.timestep()
    assert g.timestep() == dt
    B0 = {(name, ib): o.get(name, ib) for name in ["B" + s for s in AX] for ib in range(case.nb)}
    bscale = max(np.abs(v).max() for v in B0.values())
    g.op("res_solve", dt)
    its, rel = o.res_solve(dt)
    n, gits, grel = g.resist_stats()
    print(f"[res] unknowns={n} gpu its={gits} rel={grel:.2e}; oracle its={its} rel={rel:.2e}", flush=True)
    assert grel <= 1e-10 or gits == 0
    for name in ["Eres" + s for s in AX]:
        scale = max(max(np.abs(o.get(name, ib)).max() for ib in range(case.nb)), 1e-300)
        for ib in range(case.nb):
            assert np.abs(g.get(name, ib) - o.get(name, ib)).max() <= 1e-8 * scale, name
    # the change of B = -dt curl(Eres) agrees to the solve tolerance
    dscale = max(np.abs(o.get(n_, ib) - B0[(n_, ib)]).max() for n_ in ["B" + s for s in AX] for ib in range(case.nb))
    for name in ["B" + s for s in AX]:
        for ib in range(case.nb):
            d = np.abs(g.get(name, ib) - o.get(name, ib)).max()
            assert d <= 1e-7 * dscale + 1e-15 * bscale, (name, d, dscale)


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("integrator", [0, 1])
def test_steps_with_resistivity(rpair, form, integrator):
    case, g, o = rpair
    g.set_form(form)
    dg = g.step(3, integrator)
    do = o.step(3, integrator)
    np.testing.assert_allclose(dg, do, rtol=1e-9)
    for name in ["U_H", "U_Na"] + ["B" + s for s in AX]:
        num = den = 0.0
        for ib in range(case.nb):
            a, b = g.get(name, ib), o.get(name, ib)
            num += float(np.sum((a - b) ** 2))
            den += float(np.sum(b ** 2))
        assert (num / den) ** 0.5 <= 1e-9, name
    assert g.divb()[0] <= 1e-12


def test_graph_capture_refused(rpair):
    case, g, o = rpair
    with pytest.raises(MercuryError):
        g.graph_capture(0)


def test_resistive_shell_mixed_form(oracle_mod):
    """Mixed explicit stages with the FP64 correction on the resistive shell:
    rel-L2 <= 1e-5 per state variable against the FP64 oracle after 5 steps."""
    case = cases.make_resistive_shell(scale=0.06)
    print(f"[case] {case.name} mixed", flush=True)
    g = Mercury(case)
    o = oracle_mod.Oracle(case)
    try:
        g.prepare()
        o.prepare()
        g.set_form(2)
        dg = g.step(5, 0)
        do = o.step(5, 0)
        np.testing.assert_allclose(dg, do, rtol=1e-5)
        for name in ("U_H", "U_Na", "B_xi", "B_eta", "B_zeta"):
            num = den = 0.0
            for ib in range(case.nb):
                a, b = g.get(name, ib), o.get(name, ib)
                num += float(np.sum((a - b) ** 2))
                den += float(np.sum(b ** 2))
            assert (num / den) ** 0.5 <= 1e-5, name
        n, its, rel = g.resist_stats()
        assert n > 0 and its > 0 and rel <= 1e-10
    finally:
        g.close()

