"""The per-step div B monitor of the fused forms (mercury_divb_monitor): the
B-cell pass folds calc_divB's maxima while it reads the synchronised face
fluxes, so the monitor's last value equals mercury_divb on the same state bit
for bit, and its running maximum equals the maximum of the per-step values."""
import numpy as np
import pytest

from mpcns_mercury_amd import cases
from mpcns_mercury_amd.hip import Mercury

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("shape", ["shell", "cut_cube"])
def test_monitor_equals_divb_every_step(form, shape):
    case = cases.make_shell(scale=0.1) if shape == "shell" else cases.make_split_cube(n=(10, 8, 6), split=(2, 1, 1))
    print(f"[case] {case.name} form={form}", flush=True)
    g = Mercury(case)
    try:
        g.prepare()
        g.set_form(form)
        g.step(1, 0)
        assert g.divb_monitor() == (0.0, 0.0, 0.0)  # off by default
        g.set_divb_monitor(True)
        g.divb_monitor(reset=True)
        per_step = []
        for _ in range(4):
            g.step(1, 0)
            last, mx, last_abs = g.divb_monitor()
            norm, maxabs = g.divb()
            assert last == norm and last_abs == maxabs, (last, norm, last_abs, maxabs)
            per_step.append(norm)
        assert g.divb_monitor()[1] == max(per_step)
        assert g.divb_monitor(reset=True)[1] == max(per_step)
        assert g.divb_monitor() == (0.0, 0.0, 0.0)
        g.graph_capture(0)
        g.graph_launch(3)
        g.synchronize()
        last, mx, _ = g.divb_monitor()
        assert last == g.divb()[0] and mx >= last > 0.0 and mx <= 1e-12
    finally:
        g.close()
