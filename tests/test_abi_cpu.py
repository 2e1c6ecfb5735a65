"""C ABI checks that need no GPU: the HIP library loads, exports every symbol
include/mercury_hip.h declares, its host-side geometry equals the oracle's bit
for bit, and invalid calls fail with a message instead of crashing."""
import ctypes
import subprocess
import os
import re

import numpy as np
import pytest

from mpcns_mercury_amd import cases, hip
from mpcns_mercury_amd.params import MercuryPhysics

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mercury_hip.h")


def _header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mercury_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = hip.load()
    names = _header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(hip.ABI_SYMBOLS)


def test_physics_struct_layout_matches_header():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct MercuryPhysics \{(.*?)\} MercuryPhysics;", txt, flags=re.S).group(1)
    n = 0
    for line in body.split(";"):
        line = line.strip()
        if not line.startswith("double"):
            continue
        for item in line[len("double"):].split(","):
            m = re.search(r"\[(\d+)\]", item)
            n += int(m.group(1)) if m else 1
    assert ctypes.sizeof(MercuryPhysics) == 8 * n


def test_field_info_host_only():
    s, n, loc = hip.field_info("U_H")
    assert n == 5 and loc == 0
    assert hip.field_info("Jcell_from_Jedge_w")[1] == 36
    with pytest.raises(hip.MercuryError):
        hip.field_info("no_such_field")


def test_invalid_calls_fail_cleanly():
    L = hip.load()
    assert L.mercury_step(None, 1, 0, None) == -1
    assert b"context" in L.mercury_last_error()
    assert L.mercury_create(None, 0, None, 4, None, None, None) == -1


GEOM = ["Jac", "JDxi", "JDet", "JDze", "Area_xi", "Area_eta", "Area_zeta", "beta_xi", "beta_eta", "beta_zeta",
        "dl_xi", "dl_eta", "dl_zeta", "alpha_xi", "alpha_eta", "alpha_zeta", "dr_xi", "dr_eta", "dr_zeta",
        "Bcell_from_Bface_w", "Jcell_from_Jedge_w", "cell_center", "taper_xi", "taper_eta", "taper_zeta",
        "eta_local_xi", "eta_local_eta", "eta_local_zeta"]


@pytest.mark.parametrize("which", ["cube", "shell"])
def test_host_geometry_bit_exact_vs_oracle(oracle_mod, which):
    c = cases.make_periodic_cube(n=(7, 6, 5)) if which == "cube" else cases.make_shell(scale=0.08)
    o = oracle_mod.Oracle(c)
    for ib in range(c.nb):
        for name in GEOM:
            a = hip.geometry_block(c, ib, name)
            b = o.get(name, ib)
            assert np.array_equal(a, b), f"{which} block {ib} {name}: max diff {np.abs(a - b).max():.3e}"


DROPIN = os.path.join(ROOT, "include", "mercury")
DROPIN_HEADERS = ["1_grid/1_MPCNS_Grid.h", "2_topology/Topology.h", "3_field/Field.h", "4_halo/Halo.h", "8_dec/DecOps.h",
                  "8_dec/DecDiagnostics.h", "MercurySolver.h"]


@pytest.mark.parametrize("header", DROPIN_HEADERS)
def test_dropin_header_is_self_contained(header, tmp_path):
    """Each C++ drop-in header compiles on its own (g++, host only)."""
    src = tmp_path / "inc.cpp"
    src.write_text(f'#include "{header}"\nint main() {{ return 0; }}\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
                        "-I", DROPIN, "-I", "/opt/rocm/include", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_caller_links_against_the_library():
    """The C++ caller of tests/cpp (built by __graft_entry__.build) resolves every
    symbol it takes from libmercury_hip.so."""
    exe = os.path.join(ROOT, "tests", "cpp", "mercury_caller")
    if not os.path.exists(exe):
        from mpcns_mercury_amd import build
        build.build_cpp_caller()
    r = subprocess.run(["ldd", "-r", exe], capture_output=True, text=True)
    assert "libmercury_hip.so" in r.stdout and "undefined symbol" not in (r.stdout + r.stderr), r.stdout + r.stderr


def _ensemble_args(runs):
    c = cases.concat_runs(runs)
    phys = (MercuryPhysics * len(runs))(*c.meta["run_phys"])
    return c, np.array(c.meta["run_nblocks"], dtype=np.int32), phys


def test_ensemble_refuses_runs_with_other_physics():
    """An ensemble context's runs may differ in hall_coef, inflow and IMF only
    (checked before anything touches the GPU)."""
    L = hip.load()
    runs = [cases.make_ensemble_run(r, scale=0.25) for r in range(2)]
    runs[1].phys.gamma = 1.4
    c, nbl, phys = _ensemble_args(runs)
    ip = ctypes.POINTER(ctypes.c_int)
    h = ctypes.c_void_p()
    dims, bc, nbr, tr, nodes = c.dims_array(), c.bc_array(), c.nbr_array(), c.trans_array(), c.nodes_flat()
    rc = L.mercury_create_ensemble(ctypes.byref(h), 2, nbl.ctypes.data_as(ip), dims.ctypes.data_as(ip), c.g, bc.ctypes.data_as(ip),
                                   nbr.ctypes.data_as(ip), tr.ctypes.data_as(ip), nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   phys)
    assert rc == -1 and b"may differ only in hall_coef" in L.mercury_last_error()


def test_ensemble_refuses_an_interface_between_runs():
    L = hip.load()
    a = cases.make_split_cube(n=(8, 8, 8), split=(2, 1, 1), g=4)
    b = cases.make_periodic_cube(n=(8, 8, 8), g=4)
    c, nbl, phys = _ensemble_args([a, b])
    nbl[:] = (1, 2)  # declare the first run as one block: its interface now reaches into run 1
    ip = ctypes.POINTER(ctypes.c_int)
    h = ctypes.c_void_p()
    dims, bc, nbr, tr, nodes = c.dims_array(), c.bc_array(), c.nbr_array(), c.trans_array(), c.nodes_flat()
    rc = L.mercury_create_ensemble(ctypes.byref(h), 2, nbl.ctypes.data_as(ip), dims.ctypes.data_as(ip), c.g, bc.ctypes.data_as(ip),
                                   nbr.ctypes.data_as(ip), tr.ctypes.data_as(ip), nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   phys)
    assert rc == -1 and b"joins two runs" in L.mercury_last_error()
