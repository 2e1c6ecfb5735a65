"""The context on a caller's HIP stream (mercury_set_stream).

The steps, graph replays and syncs of a context move to the caller's stream:
the state after steps launched there equals the oracle's, a device-to-host
copy the caller queues on the same stream right after an asynchronous graph
replay reads the stepped state (stream order, no other synchronisation), and
the context returns to its own stream on NULL.
"""
import numpy as np
import pytest

from mpcns_mercury_amd import cases
from mpcns_mercury_amd.hip import Mercury

from hipmem import Stream

pytestmark = pytest.mark.gpu


def test_context_runs_on_the_callers_stream(oracle_mod):
    case = cases.make_periodic_cube(n=(10, 8, 6), g=4)
    print(f"[case] {case.name} on a caller's stream", flush=True)
    g = Mercury(case)
    o = oracle_mod.Oracle(case)
    s = Stream()
    try:
        own = g.stream()
        g.prepare()
        o.prepare()
        g.set_stream(s.handle)
        assert g.stream() == s.handle
        np.testing.assert_array_equal(g.step(2, 0), o.step(2, 0))
        ptr = g.field_device("U_H", 0, 0)  # waits for the stream
        g.graph_capture(0)
        g.graph_launch(3)  # asynchronous on the caller's stream
        host = np.empty(case.box(0))
        s.copy_to_host(ptr, host)
        s.synchronize()
        o.step(3, 0)
        np.testing.assert_array_equal(host, o.get("U_H", 0)[0])
        g.set_stream(None)
        assert g.stream() == own
        np.testing.assert_array_equal(g.step(1, 0), o.step(1, 0))
        for name in ("U_H", "U_Na", "B_xi", "B_eta", "B_zeta"):
            np.testing.assert_array_equal(g.get(name, 0), o.get(name, 0), err_msg=name)
    finally:
        g.close()
        s.destroy()
