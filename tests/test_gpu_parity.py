"""HIP path versus the CPU oracle on identical seeded inputs (through the C ABI).

The FP64 device path is compiled without FMA contraction and performs the
oracle's operations in the oracle's order, so the comparisons below are exact
(``assert_equal``) except where noted.  Tolerances, where used, are stated.
"""
import numpy as np
import pytest

from mpcns_mercury_amd import cases
from mpcns_mercury_amd.hip import Mercury, MercuryError

pytestmark = pytest.mark.gpu

GEOM = ["Jac", "JDxi", "JDet", "JDze", "Area_xi", "Area_eta", "Area_zeta", "beta_xi", "beta_eta", "beta_zeta",
        "dl_xi", "dl_eta", "dl_zeta", "alpha_xi", "alpha_eta", "alpha_zeta", "dr_xi", "dr_eta", "dr_zeta",
        "Bcell_from_Bface_w", "Jcell_from_Jedge_w", "cell_center", "taper_xi", "taper_eta", "taper_zeta",
        "eta_local_xi", "eta_local_eta", "eta_local_zeta"]
STATE = ["U_H", "U_Na", "B_xi", "B_eta", "B_zeta"]
DERIVED = ["B_cell", "Bind_cell", "J_xi", "J_eta", "J_zeta", "J_cell", "PV_H", "PV_Na", "U_plus"]


def small_cube(**kw):
    return cases.make_periodic_cube(n=(10, 8, 6), g=4, **kw)


def small_shell(**kw):
    return cases.make_shell(scale=0.1, **kw)


def _cmp(a, b, names, ctx, rtol=0.0):
    bad = []
    for name in names:
        for ib in range(a.case.nb):
            x, y = a.get(name, ib), b.get(name, ib)
            if rtol == 0.0:
                if not np.array_equal(x, y):
                    d = np.abs(x - y)
                    bad.append(f"{name}[{ib}] max|diff|={np.nanmax(d):.3e} n={int((d > 0).sum())}")
            else:
                scale = max(np.abs(y).max(), 1e-300)
                err = np.abs(x - y).max() / scale
                if not err <= rtol:
                    bad.append(f"{name}[{ib}] rel={err:.3e}")
    assert not bad, f"{ctx}: " + "; ".join(bad)


@pytest.fixture(params=["cube", "shell", "shell_dipole"])
def pair(request, oracle_mod):
    case = {"cube": small_cube, "shell": small_shell, "shell_dipole": lambda: small_shell(dipole=True)}[request.param]()
    print(f"[case] {case.name} dims={[b.dims for b in case.blocks]}", flush=True)
    g = Mercury(case)
    o = oracle_mod.Oracle(case)
    yield case, g, o
    g.close()


def test_geometry_bit_exact(pair):
    case, g, o = pair
    _cmp(g, o, GEOM, "geometry")


def test_prepare_derived_bit_exact(pair):
    case, g, o = pair
    g.prepare()
    o.prepare()
    _cmp(g, o, STATE + ["Badd_xi", "Badd_eta", "Badd_zeta"] + DERIVED, "prepare")


def test_timestep_bit_exact(pair):
    case, g, o = pair
    g.prepare()
    o.prepare()
    assert g.timestep() == o.timestep()


def test_fluid_rhs_bit_exact(pair):
    case, g, o = pair
    g.prepare()
    o.prepare()
    g.op("flux")
    o.op("flux")
    _cmp(g, o, ["RHS_H", "RHS_Na"], "flux")
    g.op("sources")
    o.op("sources")
    _cmp(g, o, ["RHS_H", "RHS_Na"], "sources")


@pytest.mark.parametrize("shape", ["cube", "cube_wide", "shell"])
def test_fused_rhs_bit_exact(oracle_mod, shape):
    """The fused flux + sources pass equals the oracle's Scheme_U_ followed by
    AddSourceToRHS_Fluid, on three block shapes (the wide cube has rows longer
    than a wavefront and an odd row length)."""
    case = {"cube": small_cube, "cube_wide": lambda: cases.make_periodic_cube(n=(12, 24, 30), g=4),
            "shell": lambda: cases.make_shell(scale=0.25)}[shape]()
    print(f"[case] {case.name} dims={[b.dims for b in case.blocks]}", flush=True)
    g = Mercury(case)
    o = oracle_mod.Oracle(case)
    try:
        g.prepare()
        o.prepare()
        g.set_form(1)
        g.op("rhs_cell")
        o.op("flux")
        o.op("sources")
        _cmp(g, o, ["RHS_H", "RHS_Na"], "fused rhs")
    finally:
        g.close()


TILE_SHAPES = {
    "cube": small_cube,  # every tile partial (m < 16)
    "cube_tiles": lambda: cases.make_periodic_cube(n=(6, 32, 17), g=4),  # full j tiles; a one-column k tile
    "cube_wide": lambda: cases.make_periodic_cube(n=(12, 24, 30), g=4),
    "shell": lambda: cases.make_shell(scale=0.25),
    "cube_he": lambda: cases.make_periodic_cube(n=(10, 20, 18), g=4, he=True),
}


@pytest.mark.parametrize("length", [1, 3, 16])
@pytest.mark.parametrize("shape", list(TILE_SHAPES))
def test_tiled_rhs_bit_exact(oracle_mod, shape, length):
    """The LDS-tiled flux + sources pass (16 x 16 column tiles marching `length`
    planes along i) equals the oracle's Scheme_U_ followed by
    AddSourceToRHS_Fluid in FP64, and the column kernel's result in the mixed
    form, bit for bit: on partial tiles, full tiles whose upper faces lie on the
    block edge, a one-column tile, chunks that end inside and at the block end,
    and three species."""
    case = TILE_SHAPES[shape]()
    names = ["RHS_H", "RHS_Na"] + (["RHS_He"] if case.phys.n_species >= 3 else [])
    print(f"[case] {case.name} dims={[b.dims for b in case.blocks]} length={length}", flush=True)
    g = Mercury(case)
    o = oracle_mod.Oracle(case)
    try:
        g.prepare()
        o.prepare()
        g.set_form(1)
        g.set_flux_shape(True, length)
        assert g.flux_shape() == (True, length)
        g.op("rhs_cell")
        o.op("flux")
        o.op("sources")
        _cmp(g, o, names, "tiled rhs")
        g.set_form(2)
        g.prepare()
        g.set_flux_shape(False, 4)
        g.op("rhs_cell")
        col = {n: [g.get(n, ib) for ib in range(case.nb)] for n in names}
        g.set_flux_shape(True, length)
        g.op("rhs_cell")
        for n in names:
            for ib in range(case.nb):
                np.testing.assert_array_equal(g.get(n, ib), col[n][ib], err_msg=f"mixed tiled {n}[{ib}]")
    finally:
        g.close()


@pytest.mark.parametrize("integrator", [0, 1])
def test_tiled_steps_and_graph_bit_exact(pair, integrator):
    """Steps with the tiled flux pass, then a captured step replayed from a HIP
    graph, equal the oracle."""
    case, g, o = pair
    g.prepare()
    o.prepare()
    g.set_form(1)
    g.set_flux_shape(True, 5)
    np.testing.assert_array_equal(g.step(5, integrator), o.step(5, integrator))
    _cmp(g, o, STATE + DERIVED, f"5 tiled steps integrator={integrator}")
    g.graph_capture(integrator)
    g.graph_launch(3)
    g.synchronize()
    o.step(3, integrator)
    _cmp(g, o, STATE + DERIVED, f"tiled graph replay integrator={integrator}")



def test_flux_shape_refuses_zero_length(pair):
    case, g, o = pair
    with pytest.raises(MercuryError, match="length must be >= 1"):
        g.set_flux_shape(True, 0)
    with pytest.raises(MercuryError, match="kind must be"):
        g.set_flux_shape(2, 4)


def test_induction_bit_exact(pair):
    case, g, o = pair
    g.prepare()
    o.prepare()
    g.op("induction")
    o.op("induction")
    _cmp(g, o, ["Eface_xi", "Eface_eta", "Eface_zeta", "Ehall_xi", "Ehall_eta", "Ehall_zeta", "E_xi", "E_eta", "E_zeta",
                "RHS_B_xi", "RHS_B_eta", "RHS_B_zeta"], "induction")


@pytest.mark.parametrize("integrator", [0, 1])
@pytest.mark.parametrize("form", [0, 1])
def test_steps_bit_exact(pair, integrator, form):
    case, g, o = pair
    g.prepare()
    o.prepare()
    g.set_form(form)
    dg = g.step(5, integrator)
    do = o.step(5, integrator)
    np.testing.assert_array_equal(dg, do)
    _cmp(g, o, STATE + DERIVED, f"5 steps integrator={integrator}")
    ng, _ = g.divb()
    assert ng <= 1e-12


@pytest.mark.parametrize("integrator", [0, 1])
def test_hundred_steps_bit_exact(oracle_mod, integrator):
    """100 steps of the periodic cube (the single-block configuration's
    length), fused FP64 form replayed from a graph: the state equals the
    oracle's bit for bit (relative L2 = 0 <= 1e-12) and div B stays at roundoff."""
    case = cases.make_periodic_cube(n=(12, 10, 8), g=4, integrator=integrator)
    print(f"[case] {case.name} 100 steps integrator={integrator}", flush=True)
    g = Mercury(case)
    o = oracle_mod.Oracle(case)
    try:
        g.prepare()
        o.prepare()
        g.set_form(1)
        g.graph_capture(integrator)
        g.graph_launch(100)
        g.synchronize()
        o.step(100, integrator)
        _cmp(g, o, STATE + DERIVED, "100 steps")
        assert g.divb()[0] <= 1e-12
    finally:
        g.close()


def test_autotune_keeps_results(pair):
    """The launch-shape autotuner changes no result: steps after it equal the oracle."""
    case, g, o = pair
    g.prepare()
    o.prepare()
    g.set_form(1)
    L = g.autotune()
    kind, length = g.flux_shape()
    assert length == L and L in ((4, 8, 16) if kind == 1 else (1, 2, 4, 8)) and kind in (0, 1)
    np.testing.assert_array_equal(g.step(3, 0), o.step(3, 0))
    _cmp(g, o, STATE + DERIVED, f"3 steps after autotune (kind {kind}, length {L})")


@pytest.mark.parametrize("integrator", [0, 1])
def test_graph_replay_equals_oracle(pair, integrator):
    """A captured step replayed from a HIP graph gives the oracle's state."""
    case, g, o = pair
    g.prepare()
    o.prepare()
    g.set_form(1)
    g.graph_capture(integrator)
    g.graph_launch(4)
    g.synchronize()
    o.step(4, integrator)
    _cmp(g, o, STATE + DERIVED, f"graph replay integrator={integrator}")


def _rel_l2(g, o, name, nb):
    num = den = 0.0
    for ib in range(nb):
        a, b = g.get(name, ib), o.get(name, ib)
        num += float(np.sum((a - b) ** 2))
        den += float(np.sum(b ** 2))
    return (num / den) ** 0.5 if den > 0 else num ** 0.5


def test_mixed_precision_accuracy(pair):
    """Mixed form (FP32 fluxes/EMFs/Hall/reconstruction, FP64 state and CT update)
    against the FP64 oracle after 10 steps: rel-L2 <= 1e-5 per conserved variable,
    normalised div B <= 1e-12 (the targets of the mixed form)."""
    case, g, o = pair
    g.prepare()
    o.prepare()
    g.set_form(2)
    dg = g.step(10, 0)
    do = o.step(10, 0)
    np.testing.assert_allclose(dg, do, rtol=1e-5)
    errs = {n: _rel_l2(g, o, n, case.nb) for n in STATE}
    print("[mixed] rel-L2:", errs, flush=True)
    assert all(e <= 1e-5 for e in errs.values()), errs
    assert g.divb()[0] <= 1e-12


def test_mixed_precision_accuracy_100_steps(pair):
    """The mixed form against the FP64 oracle after 100 steps (the horizon of the
    FP64 bar), replayed from a graph: rel-L2 <= 1e-5 per conserved variable,
    normalised div B <= 1e-12."""
    case, g, o = pair
    g.prepare()
    o.prepare()
    g.set_form(2)
    g.graph_capture(0)
    g.graph_launch(100)
    g.synchronize()
    o.step(100, 0)
    errs = {n: _rel_l2(g, o, n, case.nb) for n in STATE}
    print("[mixed 100] rel-L2:", errs, flush=True)
    assert all(e <= 1e-5 for e in errs.values()), errs
    assert g.divb()[0] <= 1e-12


def _runaway(case, g, o=None):
    """A cell moving at 1e10 drives the CFL time step below 1e-7.  The cell has
    periodic ghost copies; the upload is synchronised (Sync of U_cell) so that
    the state recorded before a refused step is one the step's own ghost fill
    would reproduce."""
    U = g.get("U_H", 0)
    i = case.g + 2
    U[1, i, i, i] = 1e10 * U[0, i, i, i]
    g.set("U_H", 0, U)
    g.op("sync_ucell")
    if o is not None:
        o.set("U_H", 0, U)
        o.op("sync_ucell", 0)


@pytest.mark.parametrize("path", ["step", "step_dt_out", "graph"])
@pytest.mark.parametrize("form", [0, 1, 2])
def test_dt_underflow_is_an_error(oracle_mod, form, path):
    """Compute_Timestep's abort rule (MercurySolver_Timestep.cpp:129-140), applied
    on the device before the update in every path: the refused step leaves the
    state of the last completed step, the call fails through the C ABI naming
    the step, later calls are refused until prepare(), and the oracle refuses
    the same step."""
    case = small_cube()
    print(f"[case] {case.name} dt underflow form={form} path={path}", flush=True)
    g = Mercury(case)
    o = oracle_mod.Oracle(case)
    try:
        healthy = {n: g.get(n, 0) for n in STATE}
        _runaway(case, g, o)
        g.prepare()
        g.set_form(form)
        o.prepare()
        before = {n: g.get(n, 0) for n in STATE}
        if path == "graph":
            g.graph_capture(0)
            with pytest.raises(MercuryError, match="dt underflow.*step 1 "):
                g.graph_launch(3)
        else:
            with pytest.raises(MercuryError, match="dt underflow.*step 1 "):
                g.step(3, 0, record_dt=(path == "step_dt_out"))
        for n in STATE:
            np.testing.assert_array_equal(g.get(n, 0), before[n], err_msg=f"{n} changed by a refused step")
        with pytest.raises(MercuryError, match="dt underflow"):
            g.step(1, 0)
        with pytest.raises(RuntimeError):
            o.step(1, 0)
        # a healthy state after prepare() steps again
        for n in STATE:
            g.set(n, 0, healthy[n])
        g.prepare()
        g.step(2, 0)
    finally:
        g.close()


@pytest.mark.parametrize("form", [1, 2])
def test_dt_underflow_in_an_ensemble(form):
    """An ensemble replay reports the member that hit the abort rule; that
    member keeps its last completed state while the healthy one advances."""
    from mpcns_mercury_amd.hip import Ensemble
    c0, c1 = small_cube(), small_cube(seed=2)
    print(f"[case] ensemble dt underflow form={form}", flush=True)
    g0, g1 = Mercury(c0), Mercury(c1)
    try:
        _runaway(c1, g1)
        for g in (g0, g1):
            g.prepare()
            g.set_form(form)
        b0, b1 = g0.get("U_H", 0), g1.get("U_H", 0)
        ens = Ensemble([g0, g1])
        with pytest.raises(MercuryError, match="member 1: .*dt underflow"):
            ens.launch(2)
        np.testing.assert_array_equal(g1.get("U_H", 0), b1)
        assert not np.array_equal(g0.get("U_H", 0), b0)
        ens.close()
    finally:
        g0.close()
        g1.close()


def test_stale_ensemble_is_refused():
    """A member whose form changes after the ensemble was captured makes the
    ensemble refuse to launch (checked on the host, nothing replays)."""
    from mpcns_mercury_amd.hip import Ensemble
    print("[case] stale ensemble", flush=True)
    g0, g1 = Mercury(small_cube()), Mercury(small_cube(seed=2))
    try:
        for g in (g0, g1):
            g.prepare()
        ens = Ensemble([g0, g1])
        ens.launch(1)
        g1.set_form(2)
        with pytest.raises(MercuryError, match="member 1 changed form"):
            ens.launch(1)
        ens.close()
        g0.graph_capture(0)
        g0.autotune()
        with pytest.raises(MercuryError, match="no captured step graph|stale"):
            g0.graph_launch(1)
    finally:
        g0.close()
        g1.close()



STATE3 = STATE + ["U_He"]
DERIVED3 = DERIVED + ["PV_He"]


@pytest.mark.parametrize("integrator", [0, 1])
@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("shape", ["cube_he", "shell_he"])
def test_three_species_steps_bit_exact(oracle_mod, shape, form, integrator):
    """H+, Na+ and He++ (the third species has no reference counterpart; the
    oracle's is pinned by reductions to two species): 5 steps of the plain and
    fused FP64 forms equal the oracle bit for bit, div B at roundoff."""
    case = cases.make_periodic_cube(n=(10, 8, 6), g=4, he=True, hall=True) if shape == "cube_he" else small_shell(he=True)
    print(f"[case] {case.name} three species form={form} integrator={integrator}", flush=True)
    g = Mercury(case)
    o = oracle_mod.Oracle(case)
    try:
        g.prepare()
        o.prepare()
        g.set_form(form)
        np.testing.assert_array_equal(g.step(5, integrator), o.step(5, integrator))
        _cmp(g, o, STATE3 + DERIVED3, f"three species, 5 steps form={form}")
        assert g.divb()[0] <= 1e-12
    finally:
        g.close()


def test_three_species_graph_and_mixed(oracle_mod):
    """The three-species shell replayed from a graph (fused FP64: bit-exact) and
    in the mixed form (rel-L2 <= 1e-5 per conserved variable after 20 steps)."""
    case = small_shell(he=True)
    print(f"[case] {case.name} three species graph + mixed", flush=True)
    o = oracle_mod.Oracle(case)
    o.prepare()
    o.step(20, 0)
    for form in (1, 2):
        g = Mercury(case)
        try:
            g.prepare()
            g.set_form(form)
            g.graph_capture(0)
            g.graph_launch(20)
            g.synchronize()
            if form == 1:
                _cmp(g, o, STATE3, "three species graph")
            else:
                errs = {n: _rel_l2(g, o, n, case.nb) for n in STATE3}
                print("[mixed, three species] rel-L2:", errs, flush=True)
                assert all(e <= 1e-5 for e in errs.values()), errs
            assert g.divb()[0] <= 1e-12
        finally:
            g.close()
