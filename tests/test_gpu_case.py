"""A case read back from CASE geometry files (written over two ranks, so with
inner and parallel connections) runs through the C ABI exactly as the
in-memory case it was written from."""
import numpy as np
import pytest

from mpcns_mercury_amd import cases, hip
from mpcns_mercury_amd.hip import Mercury

import case_format as cf
from test_gpu_parity import STATE

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", ["cube", "sector"])
def test_case_files_run_like_the_case(tmp_path, which):
    if which == "cube":
        c = cases.make_split_cube((12, 10, 8), (2, 2, 1))
    else:
        c = cases.make_split_sector(dims=(16, 16, 12), cuts=((8, 8), (8, 8), (12,)), scale=1.0)
    print(f"[case] {c.name} from case files", flush=True)
    cf.write_case(tmp_path, c, [0, 0, 1, 1])
    geo = hip.CaseGeometry(tmp_path)
    r = cases.case_from_geometry(geo, c)
    geo.close()
    a, b = Mercury(c), Mercury(r)
    try:
        for x in (a, b):
            x.prepare()
        np.testing.assert_array_equal(a.step(4, 0), b.step(4, 0))
        for name in STATE:
            for ib in range(c.nb):
                np.testing.assert_array_equal(a.get(name, ib), b.get(name, ib), err_msg=f"{name}[{ib}]")
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("rule", [0, 1])
def test_cubed_sphere_case_files_run_like_the_case(tmp_path, rule):
    """A cubed sphere written as CASE files (its rotated panel connections as inner
    records) and read back: the nodes and the patch list the reader builds by the
    reference's IndexTransform rule drive a patch context, singular corner edges
    included, exactly like the in-memory case."""
    def sphere():
        s = cases.make_cubed_sphere(6, 4, rule=rule)
        s.Bu = (0.05, -0.03, 0.08)
        return s
    c = sphere()
    c.A = cases.edge_potential(c, lambda x, y, z: (0.01 * y * z, 0.0 * x, -0.02 * (x * x + y * y)))
    print(f"[case] {c.name} rule={rule} from case files", flush=True)
    cf.write_case(tmp_path, c, [0] * c.nb)
    geo = hip.CaseGeometry(tmp_path)
    r = sphere()
    for b in range(r.nb):
        r.blocks[b].nodes = np.asarray(geo.nodes[b])
    bc, r.patches = geo.patch_topology()
    geo.close()
    assert (bc == r.bc_array()).all()
    r.A = cases.edge_potential(r, lambda x, y, z: (0.01 * y * z, 0.0 * x, -0.02 * (x * x + y * y)))
    a, b = Mercury(c), Mercury(r)
    try:
        for x in (a, b):
            x.set_form(1)
            x.prepare()
        np.testing.assert_array_equal(a.step(3, 0), b.step(3, 0))
        for name in STATE:
            for ib in range(c.nb):
                np.testing.assert_array_equal(a.get(name, ib), b.get(name, ib), err_msg=f"{name}[{ib}]")
    finally:
        a.close()
        b.close()
