"""ctypes front end of ``libmercury_hip.so`` (the C ABI of include/mercury_hip.h).

The HIP library is the only compute path: if it is missing or no GPU is
present, every entry point raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import MercuryPhysics

_HERE = os.path.dirname(os.path.abspath(__file__))
# MERCURY_HIP_LIB selects another build of the same ABI, e.g. the host-checked
# libmercury_hip_checked.so (mpcns_mercury_amd/build.py)
LIB_PATH = os.environ.get("MERCURY_HIP_LIB") or os.path.join(_HERE, "libmercury_hip.so")
_lib = None

# every symbol declared in include/mercury_hip.h
ABI_SYMBOLS = ("mercury_last_error", "mercury_version", "mercury_create", "mercury_create_connected", "mercury_create_patched",
               "mercury_patch_plan", "mercury_singular_edges", "mercury_create_ensemble", "mercury_run_dt", "mercury_destroy",
               "mercury_field_info",
               "mercury_block_volume", "mercury_block_shape", "mercury_field_device", "mercury_stream", "mercury_set_stream", "mercury_total_volume", "mercury_upload", "mercury_download",
               "mercury_init_bface", "mercury_init_bface_dipoles", "mercury_prepare", "mercury_step", "mercury_op", "mercury_timestep",
               "mercury_divb", "mercury_set_divb_monitor", "mercury_divb_monitor", "mercury_synchronize", "mercury_graph_capture", "mercury_graph_launch",
               "mercury_time_op", "mercury_geometry_block", "mercury_set_form", "mercury_resist_stats",
               "mercury_ensemble_create", "mercury_ensemble_launch", "mercury_ensemble_destroy", "mercury_autotune",
               "mercury_set_flux_shape", "mercury_flux_shape", "mercury_set_sync_plan", "mercury_sync_plan",
               "mercury_set_alias_classes", "mercury_alias_classes",
               "mercury_write_restart", "mercury_read_restart", "mercury_write_rundata", "mercury_read_rundata",
               "mercury_write_tecplot", "mercury_case_read", "mercury_case_destroy", "mercury_case_blocks",
               "mercury_case_block_name", "mercury_case_nodes", "mercury_case_topology", "mercury_case_interfaces",
               "mercury_case_physical", "mercury_case_patch_topology", "mercury_dec_d0", "mercury_dec_d1", "mercury_dec_d2", "mercury_dec_faraday_update",
               "mercury_dec_max_abs_div_curl", "mercury_dec_max_abs_curl_grad", "mercury_dec_max_abs_div_curl_noalloc")


class MercuryError(RuntimeError):
    pass


def load():
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MercuryError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or python -m mpcns_mercury_amd.build)")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp, cp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), ctypes.c_char_p
    L.mercury_last_error.restype = ctypes.c_char_p
    L.mercury_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ip, ctypes.c_int, ip, dp, ctypes.POINTER(MercuryPhysics)]
    L.mercury_create_connected.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ip, ctypes.c_int, ip, ip, ip, dp,
                                           ctypes.POINTER(MercuryPhysics)]
    L.mercury_create_patched.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ip, ctypes.c_int, ip, ctypes.c_int, ip, ctypes.c_int, dp,
                                         ctypes.POINTER(MercuryPhysics)]
    L.mercury_patch_plan.argtypes = [ctypes.c_int, ip, ctypes.c_int, ip, ctypes.c_int, ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong), ip]
    L.mercury_singular_edges.argtypes = [ctypes.c_int, ip, ctypes.c_int, ip, ctypes.c_int, ip, ctypes.c_int, dp,
                                         ctypes.POINTER(MercuryPhysics), ctypes.c_int, ip, ip, dp]
    L.mercury_destroy.argtypes = [vp]
    L.mercury_create_ensemble.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ip, ip, ctypes.c_int, ip, ip, ip, dp,
                                          ctypes.POINTER(MercuryPhysics)]
    L.mercury_run_dt.argtypes = [vp, ip, dp]
    L.mercury_write_restart.argtypes = [vp, cp, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    L.mercury_read_restart.argtypes = [vp, cp, ip, dp]
    L.mercury_write_rundata.argtypes = [cp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, ctypes.c_double]
    L.mercury_read_rundata.argtypes = [cp, ip, dp, dp, ctypes.c_int, ip, dp, dp]
    L.mercury_write_tecplot.argtypes = [vp, cp, ctypes.c_double, ctypes.POINTER(cp), ctypes.c_int, ctypes.POINTER(cp)]
    L.mercury_case_read.argtypes = [ctypes.POINTER(vp), cp, ctypes.c_int, ctypes.c_int]
    L.mercury_case_destroy.argtypes = [vp]
    L.mercury_case_destroy.restype = None
    L.mercury_case_blocks.argtypes = [vp, ip, ip]
    L.mercury_case_block_name.argtypes = [vp, ctypes.c_int, cp, ctypes.c_int]
    L.mercury_case_nodes.argtypes = [vp, dp]
    L.mercury_case_topology.argtypes = [vp, ip, ip, ip]
    L.mercury_case_interfaces.argtypes = [vp, ip, ip]
    L.mercury_case_physical.argtypes = [vp, ip, ip]
    L.mercury_field_info.argtypes = [cp, ip, ip, ip]
    L.mercury_block_volume.restype = ctypes.c_longlong
    L.mercury_block_volume.argtypes = [vp, ctypes.c_int]
    L.mercury_total_volume.restype = ctypes.c_longlong
    L.mercury_total_volume.argtypes = [vp]
    L.mercury_upload.argtypes = [vp, cp, ctypes.c_int, ctypes.c_int, dp]
    L.mercury_block_shape.argtypes = [vp, ctypes.c_int, ip, ip]
    L.mercury_field_device.argtypes = [vp, cp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    L.mercury_stream.argtypes = [vp, ctypes.POINTER(vp)]
    L.mercury_set_stream.argtypes = [vp, vp]
    L.mercury_download.argtypes = [vp, cp, ctypes.c_int, ctypes.c_int, dp]
    L.mercury_init_bface.argtypes = [vp, dp, ctypes.POINTER(dp)]
    L.mercury_init_bface_dipoles.argtypes = [vp, dp, ctypes.POINTER(dp), ctypes.c_int, dp]
    L.mercury_prepare.argtypes = [vp]
    L.mercury_step.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    L.mercury_op.argtypes = [vp, cp, ctypes.c_double]
    L.mercury_timestep.argtypes = [vp, dp]
    L.mercury_divb.argtypes = [vp, dp, dp]
    L.mercury_synchronize.argtypes = [vp]
    L.mercury_graph_capture.argtypes = [vp, ctypes.c_int]
    L.mercury_set_form.argtypes = [vp, ctypes.c_int]
    L.mercury_graph_launch.argtypes = [vp, ctypes.c_int]
    L.mercury_time_op.argtypes = [vp, cp, ctypes.c_int, dp]
    L.mercury_resist_stats.argtypes = [vp, ip, ip, dp]
    L.mercury_ensemble_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    L.mercury_ensemble_launch.argtypes = [vp, ctypes.c_int]
    L.mercury_ensemble_destroy.argtypes = [vp]
    L.mercury_autotune.argtypes = [vp, ip]
    L.mercury_set_flux_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.mercury_divb_monitor.argtypes = [vp, dp, dp, dp, ctypes.c_int]
    L.mercury_set_divb_monitor.argtypes = [vp, ctypes.c_int]
    llp = ctypes.POINTER(ctypes.c_longlong)
    L.mercury_set_sync_plan.argtypes = [vp, cp, ctypes.c_int, ctypes.c_longlong, ip]
    L.mercury_sync_plan.argtypes = [vp, cp, ctypes.c_int, ctypes.c_longlong, llp, ip]
    L.mercury_set_alias_classes.argtypes = [vp, ctypes.c_int, ip, ip, ip, ip]
    L.mercury_alias_classes.argtypes = [vp, ctypes.c_int, ctypes.c_longlong, ip, llp, ip, ip, ip, ip]
    L.mercury_flux_shape.argtypes = [vp, ip, ip]
    L.mercury_geometry_block.argtypes = [ip, ctypes.c_int, dp, ctypes.POINTER(MercuryPhysics), cp, ctypes.c_int, dp]
    # DEC operators on caller device arrays: pointers travel as void *
    pp = ctypes.POINTER(vp)
    L.mercury_dec_d0.argtypes = [ip, ctypes.c_int, vp, pp, vp]
    L.mercury_dec_d1.argtypes = [ip, ctypes.c_int, pp, pp, vp]
    L.mercury_dec_d2.argtypes = [ip, ctypes.c_int, pp, vp, vp]
    L.mercury_dec_faraday_update.argtypes = [ip, ctypes.c_int, pp, pp, ctypes.c_double, vp]
    L.mercury_dec_max_abs_div_curl.argtypes = [ip, ctypes.c_int, pp, vp, vp]
    L.mercury_dec_max_abs_curl_grad.argtypes = [ip, ctypes.c_int, vp, vp, vp]
    L.mercury_dec_max_abs_div_curl_noalloc.argtypes = [ip, ctypes.c_int, pp, vp, vp]
    _lib = L
    return L


def _chk(rc):
    if rc != 0:
        raise MercuryError(load().mercury_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def write_rundata(path, step, time, dt, residual_ref=(), residual_max=0.0):
    """MPCNSRUN v1 run-data file (monitors empty)."""
    ref = np.ascontiguousarray(residual_ref, dtype=np.float64)
    _chk(load().mercury_write_rundata(str(path).encode(), int(step), float(time), float(dt), len(ref), _dp(ref),
                                      float(residual_max)))


def read_rundata(path, max_ref=64):
    """(step, time, dt, residual_ref, residual_max) of an MPCNSRUN v1 file."""
    st, tm, dt, n, rmax = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
    ref = np.zeros(max_ref)
    _chk(load().mercury_read_rundata(str(path).encode(), ctypes.byref(st), ctypes.byref(tm), ctypes.byref(dt), max_ref,
                                     ctypes.byref(n), _dp(ref), ctypes.byref(rmax)))
    return st.value, tm.value, dt.value, ref[:n.value].copy(), rmax.value


def field_info(name):
    s, n, l = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _chk(load().mercury_field_info(name.encode(), ctypes.byref(s), ctypes.byref(n), ctypes.byref(l)))
    return s.value, n.value, l.value


def geometry_block(case, ib, name):
    """Host-side static geometry of one block (no GPU involved)."""
    _, nc, _ = field_info(name)
    b = case.blocks[ib]
    dims = np.array(b.dims, dtype=np.int32)
    nodes = np.ascontiguousarray(b.nodes, dtype=np.float64).reshape(-1)
    shape = case.box(ib)
    out = np.empty((nc,) + shape)
    for c in range(nc):
        buf = np.empty(shape)
        _chk(load().mercury_geometry_block(dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), case.g, _dp(nodes),
                                           ctypes.byref(case.phys), name.encode(), c, _dp(buf)))
        out[c] = buf
    return out


PATCH_KINDS = {"component": 0, "face": 1, "edge": 2}


def patch_plan(case, kind: str, level: int, stage: int) -> np.ndarray:
    """The copies a patch-connected context derives for one kind of field group
    (host only, no GPU): rows (dst axis, block, i, j, k, src axis, block, i, j, k, sign)."""
    L = load()
    ip = ctypes.POINTER(ctypes.c_int)
    dims, bc = case.dims_array(), case.bc_array()
    pt = np.ascontiguousarray(case.patches, dtype=np.int32).reshape(-1)
    n = ctypes.c_longlong(0)
    args = (case.nb, dims.ctypes.data_as(ip), case.g, bc.ctypes.data_as(ip), len(pt) // 18, pt.ctypes.data_as(ip), int(case.patch_rule),
            PATCH_KINDS[kind], int(level), int(stage))
    _chk(L.mercury_patch_plan(*args, 0, ctypes.byref(n), None))
    out = np.zeros((n.value, 11), dtype=np.int32)
    _chk(L.mercury_patch_plan(*args, n.value, ctypes.byref(n), out.ctypes.data_as(ip)))
    return out


def singular_edges(case) -> list:
    """The singular edges a patch-connected context assembles specially (host
    only, no GPU): per entry the owner DOF (block, i, j, k, axis), its
    orientation sign, the incident cell and face counts, the inverse Hodge
    scale and the mean edge vector."""
    L = load()
    ip = ctypes.POINTER(ctypes.c_int)
    dims, bc = case.dims_array(), case.bc_array()
    pt = np.ascontiguousarray(case.patches, dtype=np.int32).reshape(-1)
    nodes = case.nodes_flat()
    n = ctypes.c_int(0)
    args = (case.nb, dims.ctypes.data_as(ip), case.g, bc.ctypes.data_as(ip), len(pt) // 18, pt.ctypes.data_as(ip), int(case.patch_rule),
            _dp(nodes), ctypes.byref(case.phys))
    _chk(L.mercury_singular_edges(*args, 0, ctypes.byref(n), None, None))
    iv = np.zeros((max(n.value, 1), 8), dtype=np.int32)
    dv = np.zeros((max(n.value, 1), 4))
    _chk(L.mercury_singular_edges(*args, n.value, ctypes.byref(n), iv.ctypes.data_as(ip), _dp(dv)))
    return [dict(owner=tuple(int(x) for x in iv[t, :5]), sign=int(iv[t, 5]), cells=int(iv[t, 6]), faces=int(iv[t, 7]),
                 inverse_hodge=float(dv[t, 0]), edge_vector=dv[t, 1:4].copy()) for t in range(n.value)]


class Mercury:
    """All blocks of one case resident on the GPU."""

    def __init__(self, case):
        L = load()
        self.case = case
        self.phys = case.phys
        self._h = ctypes.c_void_p()
        dims = case.dims_array()
        bc = case.bc_array()
        nodes = case.nodes_flat()
        ip = ctypes.POINTER(ctypes.c_int)
        if "run_nblocks" in case.meta:
            # an ensemble of independent runs in one context (cases.concat_runs)
            runs = case.meta["run_nblocks"]
            nb_arr = np.array(runs, dtype=np.int32)
            phys = (MercuryPhysics * len(runs))(*case.meta["run_phys"])
            nbr, trans = case.nbr_array(), case.trans_array()
            _chk(L.mercury_create_ensemble(ctypes.byref(self._h), len(runs), nb_arr.ctypes.data_as(ip), dims.ctypes.data_as(ip),
                                           case.g, bc.ctypes.data_as(ip), nbr.ctypes.data_as(ip), trans.ctypes.data_as(ip),
                                           _dp(nodes), phys))
        elif getattr(case, "patches", None) is not None:
            pt = np.ascontiguousarray(case.patches, dtype=np.int32).reshape(-1)
            _chk(L.mercury_create_patched(ctypes.byref(self._h), case.nb, dims.ctypes.data_as(ip), case.g, bc.ctypes.data_as(ip),
                                          len(pt) // 18, pt.ctypes.data_as(ip), int(case.patch_rule), _dp(nodes),
                                          ctypes.byref(self.phys)))
        elif case.has_interfaces():
            nbr, trans = case.nbr_array(), case.trans_array()
            _chk(L.mercury_create_connected(ctypes.byref(self._h), case.nb, dims.ctypes.data_as(ip), case.g, bc.ctypes.data_as(ip),
                                            nbr.ctypes.data_as(ip), trans.ctypes.data_as(ip), _dp(nodes),
                                            ctypes.byref(self.phys)))
        else:
            _chk(L.mercury_create(ctypes.byref(self._h), case.nb, dims.ctypes.data_as(ip), case.g, bc.ctypes.data_as(ip),
                                  _dp(nodes), ctypes.byref(self.phys)))
        for ib, fl in enumerate(case.fields):
            for name, arr in fl.items():
                self.set(name, ib, arr)
        Bu = np.array(case.Bu, dtype=np.float64)
        dip = case.dipole_array()
        ptrs = None
        if case.A is not None:
            flat = [np.ascontiguousarray(case.A[ib][a], dtype=np.float64) for ib in range(case.nb) for a in range(3)]
            self._A = flat
            ptrs = (ctypes.POINTER(ctypes.c_double) * len(flat))(*[_dp(x) for x in flat])
        _chk(L.mercury_init_bface_dipoles(self._h, _dp(Bu), ptrs, len(dip), _dp(dip)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load().mercury_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get(self, name, ib=0):
        _, nc, _ = field_info(name)
        shape = self.case.box(ib)
        out = np.empty((nc,) + shape)
        for c in range(nc):
            buf = np.empty(shape)
            _chk(load().mercury_download(self._h, name.encode(), ib, c, _dp(buf)))
            out[c] = buf
        return out

    def set(self, name, ib, arr):
        shape = self.case.box(ib)
        arr = np.asarray(arr, dtype=np.float64).reshape((-1,) + shape)
        for c in range(arr.shape[0]):
            buf = np.ascontiguousarray(arr[c])
            _chk(load().mercury_upload(self._h, name.encode(), ib, c, _dp(buf)))

    def prepare(self):
        _chk(load().mercury_prepare(self._h))

    def step(self, n=1, integrator=0, record_dt=True):
        if record_dt:
            dts = np.zeros(max(n, 1))
            _chk(load().mercury_step(self._h, n, integrator, _dp(dts)))
            return dts[:n]
        _chk(load().mercury_step(self._h, n, integrator, None))
        return None

    def op(self, name, arg=0.0):
        _chk(load().mercury_op(self._h, name.encode(), float(arg)))

    def run_dts(self):
        """The dt each run of an ensemble context committed for its last step."""
        n = ctypes.c_int(0)
        _chk(load().mercury_run_dt(self._h, ctypes.byref(n), None))
        out = np.zeros(n.value)
        _chk(load().mercury_run_dt(self._h, ctypes.byref(n), _dp(out)))
        return out

    def timestep(self):
        v = ctypes.c_double()
        _chk(load().mercury_timestep(self._h, ctypes.byref(v)))
        return v.value

    def divb(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _chk(load().mercury_divb(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_divb_monitor(self, on=True):
        """Switch the per-step div B monitor of the fused forms on or off."""
        _chk(load().mercury_set_divb_monitor(self._h, 1 if on else 0))

    def divb_monitor(self, reset=False):
        """(last normalised div B of the derived pass, its max since the last reset,
        last max |d2 B|), from the fused forms' B-cell pass; reset clears them."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _chk(load().mercury_divb_monitor(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), 1 if reset else 0))
        return a.value, b.value, c.value

    def synchronize(self):
        _chk(load().mercury_synchronize(self._h))

    def set_form(self, form):
        """0: one kernel per reference function; 1: fused FP64 passes."""
        _chk(load().mercury_set_form(self._h, int(form)))

    def graph_capture(self, integrator=0):
        _chk(load().mercury_graph_capture(self._h, integrator))

    def graph_launch(self, n=1):
        _chk(load().mercury_graph_launch(self._h, n))

    def write_restart(self, path, step=0, time=0.0, with_jedge=False):
        """MPCNSRST v1 restart of U_H, U_Na, B_xi/eta/zeta (and the J edges)."""
        _chk(load().mercury_write_restart(self._h, str(path).encode(), int(step), float(time), 1 if with_jedge else 0))

    def read_restart(self, path):
        """Load a restart file into the device fields; returns (step, time).
        Call prepare() afterwards."""
        st, tm = ctypes.c_int(), ctypes.c_double()
        _chk(load().mercury_read_restart(self._h, str(path).encode(), ctypes.byref(st), ctypes.byref(tm)))
        return st.value, tm.value

    def write_tecplot(self, path, time=0.0, fields=None, block_names=None):
        """Tecplot binary of the listed cell/node fields at the nodes (default: the
        Mercury list PV_H, PV_Na, B_cell, Bind_cell, Na, J_cell)."""
        def _strs(xs):
            if xs is None:
                return None, 0
            arr = (ctypes.c_char_p * len(xs))(*[str(x).encode() for x in xs])
            return arr, len(xs)
        fl, nf = _strs(fields)
        bn, nb = _strs(block_names)
        if bn is not None and nb != self.case.nb:
            raise ValueError("block_names needs one name per block")
        _chk(load().mercury_write_tecplot(self._h, str(path).encode(), float(time), fl, nf, bn))

    def resist_stats(self):
        """(unknown edges, iterations of the last solve, its relative residual)."""
        n, it, rel = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _chk(load().mercury_resist_stats(self._h, ctypes.byref(n), ctypes.byref(it), ctypes.byref(rel)))
        return n.value, it.value, rel.value

    def autotune(self):
        """Pick the fastest launch shape of the flux pass; returns its length
        (cells per thread, or planes per tile: see flux_shape)."""
        v = ctypes.c_int()
        _chk(load().mercury_autotune(self._h, ctypes.byref(v)))
        return v.value

    def stream(self):
        """The context's HIP stream handle (an integer hipStream_t)."""
        v = ctypes.c_void_p()
        _chk(load().mercury_stream(self._h, ctypes.byref(v)))
        return v.value

    def set_stream(self, handle):
        """Run the context's work on a caller's HIP stream (None: its own)."""
        _chk(load().mercury_set_stream(self._h, ctypes.c_void_p(handle) if handle else None))

    def field_device(self, name, block=0, comp=0):
        """Device address of one component of a field on one block."""
        v = ctypes.c_void_p()
        _chk(load().mercury_field_device(self._h, name.encode(), block, comp, ctypes.byref(v)))
        return v.value

    def sync_plan(self, group, stage):
        """The (n, 11) int32 entries of one stage (0 halo, 1 owner) of a sync group."""
        n = ctypes.c_longlong()
        _chk(load().mercury_sync_plan(self._h, group.encode(), stage, 0, ctypes.byref(n), None))
        out = np.zeros((max(n.value, 1), 11), dtype=np.int32)
        _chk(load().mercury_sync_plan(self._h, group.encode(), stage, n.value, ctypes.byref(n),
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out[:n.value]

    def set_sync_plan(self, group, stage, entries):
        e = np.ascontiguousarray(entries, dtype=np.int32).reshape(-1, 11)
        _chk(load().mercury_set_sync_plan(self._h, group.encode(), stage, e.shape[0],
                                          e.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def alias_classes(self):
        """(start, axis, members (m, 5), owners (c, 5)) of the alias-mean classes."""
        nc, nm = ctypes.c_int(), ctypes.c_longlong()
        _chk(load().mercury_alias_classes(self._h, 0, 0, ctypes.byref(nc), ctypes.byref(nm), None, None, None, None))
        ip = ctypes.POINTER(ctypes.c_int)
        start = np.zeros(nc.value + 1, np.int32)
        axis = np.zeros(max(nc.value, 1), np.int32)
        mem = np.zeros((max(nm.value, 1), 5), np.int32)
        own = np.zeros((max(nc.value, 1), 5), np.int32)
        _chk(load().mercury_alias_classes(self._h, nc.value, nm.value, ctypes.byref(nc), ctypes.byref(nm), start.ctypes.data_as(ip),
                                          axis.ctypes.data_as(ip), mem.ctypes.data_as(ip), own.ctypes.data_as(ip)))
        return start, axis[:nc.value], mem[:nm.value], own[:nc.value]

    def set_alias_classes(self, start, axis, members, owners):
        ip = ctypes.POINTER(ctypes.c_int)
        arrs = [np.ascontiguousarray(a, dtype=np.int32) for a in (start, axis, members, owners)]
        _chk(load().mercury_set_alias_classes(self._h, len(arrs[1]), *[a.ctypes.data_as(ip) for a in arrs]))

    def set_flux_shape(self, kind, length):
        """Run the flux pass as the column kernel (kind 0 or False: length cells
        per thread along i) or the LDS-tiled kernel (1 or True: length planes
        per tile)."""
        _chk(load().mercury_set_flux_shape(self._h, int(kind), int(length)))

    def flux_shape(self):
        """(kind, length) of the flux pass's launch shape."""
        t, n = ctypes.c_int(), ctypes.c_int()
        _chk(load().mercury_flux_shape(self._h, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def time_op(self, op, reps=10):
        v = ctypes.c_double()
        _chk(load().mercury_time_op(self._h, op.encode(), reps, ctypes.byref(v)))
        return v.value


class Ensemble:
    """Independent contexts stepped together by one captured graph."""

    def __init__(self, members, integrator=0):
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        h = ctypes.c_void_p()
        _chk(load().mercury_ensemble_create(arr, len(self.members), int(integrator), ctypes.byref(h)))
        self._h = h

    def launch(self, nsteps=1):
        _chk(load().mercury_ensemble_launch(self._h, int(nsteps)))

    def close(self):
        if getattr(self, "_h", None):
            load().mercury_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CaseGeometry:
    """A CASE geometry directory read by the library's C++ reader
    (mercury_case_read): blocks of every rank in rank-major order, their
    physical and connection patches, and the per-face topology for a solver."""

    def __init__(self, geometry_dir, nrank=0, binary=True):
        L = load()
        h = ctypes.c_void_p()
        _chk(L.mercury_case_read(ctypes.byref(h), str(geometry_dir).encode(), int(nrank), 1 if binary else 0))
        self._h = h
        nb = ctypes.c_int()
        _chk(L.mercury_case_blocks(self._h, ctypes.byref(nb), None))
        self.nb = nb.value
        dims = np.zeros(3 * self.nb, dtype=np.int32)
        _chk(L.mercury_case_blocks(self._h, ctypes.byref(nb), dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        self.dims = [tuple(int(x) for x in dims[3 * b:3 * b + 3]) for b in range(self.nb)]
        self.names = []
        buf = ctypes.create_string_buffer(256)
        for b in range(self.nb):
            _chk(L.mercury_case_block_name(self._h, b, buf, 256))
            self.names.append(buf.value.decode())
        flat = np.zeros(sum(3 * (m[0] + 1) * (m[1] + 1) * (m[2] + 1) for m in self.dims))
        _chk(L.mercury_case_nodes(self._h, flat.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        self.nodes, o = [], 0
        for m in self.dims:
            n = 3 * (m[0] + 1) * (m[1] + 1) * (m[2] + 1)
            self.nodes.append(flat[o:o + n].reshape(3, m[0] + 1, m[1] + 1, m[2] + 1).copy())
            o += n

    def _records(self, fn, width):
        n = ctypes.c_int()
        _chk(fn(self._h, ctypes.byref(n), None))
        rec = np.zeros(max(n.value, 1) * width, dtype=np.int32)
        _chk(fn(self._h, ctypes.byref(n), rec.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return rec[:n.value * width].reshape(n.value, width)

    def interfaces(self):
        """Connection patches: dicts with kind ('inner' / 'parallel'), periodic, block,
        nb_block, direction, nb_direction, lo, hi, nb_lo, nb_hi, perm, sign, offset."""
        out = []
        for r in self._records(load().mercury_case_interfaces, 27):
            r = [int(x) for x in r]
            out.append({"kind": "parallel" if r[0] else "inner", "periodic": bool(r[1]), "block": r[2], "nb_block": r[3],
                        "direction": r[4], "nb_direction": r[5], "lo": tuple(r[6:9]), "hi": tuple(r[9:12]),
                        "nb_lo": tuple(r[12:15]), "nb_hi": tuple(r[15:18]), "perm": tuple(r[18:21]),
                        "sign": tuple(r[21:24]), "offset": tuple(r[24:27])})
        return out

    def patch_topology(self):
        """(bc, patches (n, 18)) for mercury_create_patched: any connection patches."""
        L = load()
        ip = ctypes.POINTER(ctypes.c_int)
        bc = np.zeros(6 * self.nb, dtype=np.int32)
        n = ctypes.c_int(0)
        _chk(L.mercury_case_patch_topology(self._h, bc.ctypes.data_as(ip), 0, ctypes.byref(n), None))
        pt = np.zeros((max(n.value, 1), 18), dtype=np.int32)
        _chk(L.mercury_case_patch_topology(self._h, bc.ctypes.data_as(ip), n.value, ctypes.byref(n), pt.ctypes.data_as(ip)))
        return bc, pt[:n.value]

    def physical(self):
        """Physical patches: dicts with block, direction, bc_id, lo, hi (half-open node box)."""
        return [{"block": int(r[0]), "direction": int(r[1]), "bc_id": int(r[2]), "lo": tuple(int(x) for x in r[3:6]),
                 "hi": tuple(int(x) for x in r[6:9])} for r in self._records(load().mercury_case_physical, 9)]

    def topology(self):
        """(bc, nbr, transform) per block face for mercury_create_connected; raises
        MercuryError for faces that are not one whole-face patch."""
        bc = np.zeros(6 * self.nb, dtype=np.int32)
        nbr = np.zeros(6 * self.nb, dtype=np.int32)
        tr = np.zeros(54 * self.nb, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        _chk(load().mercury_case_topology(self._h, bc.ctypes.data_as(ip), nbr.ctypes.data_as(ip), tr.ctypes.data_as(ip)))
        return bc, nbr, tr

    def close(self):
        if getattr(self, "_h", None):
            load().mercury_case_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
