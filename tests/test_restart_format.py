"""Run-data files of the library against the restatement of the reference's
MPCNSRUN format (CPU only: the run-data entry points touch no device).

Restart files need device fields; their tests are in test_gpu_restart.py."""
import numpy as np
import pytest

from mpcns_mercury_amd import hip

import restart_format as rf


@pytest.fixture(scope="module")
def lib():
    return hip.load()


def test_rundata_written_by_the_library_parses(tmp_path, lib):
    p = tmp_path / "RunData.bin"
    ref = np.array([1.0, 0.5, 3.25e-7, -2.0])
    hip.write_rundata(p, 1234, 0.75, 1.5e-4, ref, 8.5e-3)
    step, time, dt, r, rmax, md, mi = rf.read_rundata(p)
    assert (step, time, dt, rmax) == (1234, 0.75, 1.5e-4, 8.5e-3)
    np.testing.assert_array_equal(r, ref)
    assert md == {} and mi == {}


def test_rundata_reader_takes_reference_files(tmp_path, lib):
    p = tmp_path / "RunData.bin"
    ref = np.linspace(0.1, 1.3, 13)  # the Mercury solver keeps 13 residual references
    rf.write_rundata(p, 77, 12.5, 2e-3, ref, 0.125, md={"Fluid/H/rho_min": 0.5, "MHD/Bmax": 2.0}, mi={"nfail": 3})
    step, time, dt, r, rmax = hip.read_rundata(p)
    assert (step, time, dt, rmax) == (77, 12.5, 2e-3, 0.125)
    np.testing.assert_array_equal(r, ref)
    with pytest.raises(hip.MercuryError):
        hip.read_rundata(p, max_ref=4)


def test_rundata_rejects_other_files(tmp_path, lib):
    p = tmp_path / "bad.bin"
    p.write_bytes(b"MPCNSRST" + b"\0" * 16)
    with pytest.raises(hip.MercuryError):
        hip.read_rundata(p)


def test_restart_restatement_round_trip(tmp_path):
    """The restatement itself: arrays survive write -> read in (i, j, k, m) order."""
    rng = np.random.default_rng(2)
    a = rng.standard_normal((5, 3, 4, 2))
    p = tmp_path / "r.bin"
    rf.write_restart(p, 9, 0.5, 1, [("U_H", 0, 5, 1, [((-1, -1, -1), (2, 3, 1), a)])])
    raw = p.read_bytes()
    # header 8 + 4*2 + 8 + 4*2, then the name and meta, then lo/hi/active, then values i, j, k, m
    off = 8 + 4 + 4 + 8 + 4 + 4 + (4 + 3) + 12 + 24 + 4
    first = np.frombuffer(raw, "<f8", count=6, offset=off)
    np.testing.assert_array_equal(first, [a[0, 0, 0, 0], a[1, 0, 0, 0], a[2, 0, 0, 0], a[3, 0, 0, 0], a[4, 0, 0, 0], a[0, 0, 0, 1]])
    step, time, nb, f = rf.read_restart(p)
    assert (step, time, nb) == (9, 0.5, 1)
    np.testing.assert_array_equal(f["U_H"][3][0][2], a)
