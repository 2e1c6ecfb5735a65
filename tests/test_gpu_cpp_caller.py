"""A C++ host program written against the drop-in headers (include/mercury:
Grid, TOPO::Topology, Field, FieldBlock, Halo, MercurySolver, DEC) steps cases
on the GPU; its state equals the CPU oracle's bit for bit, and its DEC chain
diagnostics on the solver's own fields equal the oracle's."""
import os
import subprocess

import numpy as np
import pytest

import cpp_case_io
import oracle as orc
from mpcns_mercury_amd import cases

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "cpp", "mercury_caller")


def _patched_sample(tmp_path):
    """The reference's 8-block sample geometry as patches (position rule)."""
    import tarfile
    from mpcns_mercury_amd import hip
    with tarfile.open(os.path.join(ROOT, "tests", "golden", "z0_null_case_geometry.tar.gz")) as t:
        t.extractall(tmp_path, members=[m for m in t.getmembers() if "Zone.Identifier" not in m.name])
    geo = hip.CaseGeometry(tmp_path / "geometry_3")
    try:
        return cases.cube_on_geometry(geo, cases.cartesian_periodic_cube(64, g=4), rule=1)
    finally:
        geo.close()


def _cubed_sphere():
    """Six gnomonic panels glued by rotated patches (position rule), with a
    uniform field plus a uniform current: the corner lines are singular edges."""
    c = cases.make_cubed_sphere(8, 6, rule=1)
    c.Bu = (0.05, -0.03, 0.08)
    c.A = cases.edge_potential(c, lambda x, y, z: (0.01 * y * z, 0.0 * x, -0.02 * (x * x + y * y)))
    return c


@pytest.mark.parametrize("which", ["cube", "shell_dipole", "patched_sample", "cubed_sphere"])
def test_cpp_caller_steps_like_the_oracle(oracle_mod, tmp_path, which):
    if not os.path.exists(CALLER):
        pytest.fail(f"{CALLER} is missing: run __graft_entry__.build()")
    if which == "cube":
        case = cases.make_periodic_cube(n=(10, 8, 6), g=4)
    elif which == "shell_dipole":
        case = cases.make_shell(scale=0.1, dipole=True)
    elif which == "patched_sample":
        case = _patched_sample(tmp_path)
    else:
        case = _cubed_sphere()
    nsteps = 4
    print(f"[case] C++ caller {case.name} {nsteps} steps", flush=True)
    o = oracle_mod.Oracle(case)
    # the caller uploads the oracle's initial state (B from the case's vector potential)
    fields = [dict(case.fields[ib]) for ib in range(case.nb)]
    for ib in range(case.nb):
        for s in ("B_xi", "B_eta", "B_zeta"):
            fields[ib][s] = o.get(s, ib)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    cpp_case_io.write_input(inp, case, fields)
    r = subprocess.run([CALLER, str(inp), str(out), str(nsteps), "0"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    # the registry the drop-in METRIC::SingularEdgeRegistry reads: the eight
    # corner lines of the cubed sphere, one entry per radial edge
    nsing = 48 if which == "cubed_sphere" else 0
    assert f"[SingularEdgeRegistry] physical singular edges={nsing}" in r.stdout
    dts, divb, div_curl, curl_grad, state = cpp_case_io.read_output(out, case, nsteps)
    o.prepare()
    np.testing.assert_array_equal(dts, o.step(nsteps, 0))
    for name, _ in cpp_case_io.STATE:
        for ib in range(case.nb):
            np.testing.assert_array_equal(state[name][ib], o.get(name, ib), err_msg=f"{name}[{ib}]")
    assert divb == o.divb()[0] and divb <= 1e-12
    ref = max(orc.dec_max_abs_div_curl(b.dims, case.g, [o.get(s, ib)[0] for s in ("E_xi", "E_eta", "E_zeta")])
              for ib, b in enumerate(case.blocks))
    assert div_curl == ref
    assert curl_grad == 0.0
