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
