#!/usr/bin/env python3
"""Benchmark of the Mercury explicit step on one MI355X.

Workload (default): two-fluid Hall MHD -- H+ and Na+ plus
electron pressure on six spherical-shell sector blocks of 96x64x64, 128x48x64,
64x96x80, 80x80x48, 112x56x72 and 72x72x96 cells (2 534 400 cells), r in [1,4],
inflow u = (-1,0,0), IMF (0,0,-0.2), conducting inner sphere, far field
elsewhere, seeded 1e-3 density noise.  One "step" is one explicit step of every
block (forward Euler, the reference integrator), replayed from a captured HIP
graph with all inputs resident in HBM.

Prints one JSON line:  value = cell-updates/s of the whole job in the fused FP64
form (bit-identical to the CPU restatement); the plain FP64 and the mixed forms
are reported beside it.  roofline = the dominant kernel's algorithmic HBM bytes /
its average duration (HIP events on the solver stream); cpu_baseline = the CPU
restatement (OpenMP threads) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mpcns_mercury_amd import cases  # noqa: E402

HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) HBM bytes per unit of each kernel of the FP64 plain
# path: every input array read once, every output written once, neighbours
# served from cache.  Units: owned cells (cell kernels) or node-box points
# (edge/face kernels, one point carries up to 3 edges or faces).
ALGO_BYTES = {
    # fused passes
    # flux + sources: U_s density (2) + PV_s (10) + JD (9) + Jac + U+ (3) + B_cell (3) + J_cell (3) + Photo
    # + centre (3) read, RHS_s (10) written
    "rhs_cell": 8 * (2 + 10 + 9 + 1 + 3 + 3 + 3 + 1 + 3 + 10),
    # B (3) + Badd (3) + U_con (3) + PV p (2) read, tangential Eface (6) + node electron pressure (1) written
    "eface_u": 8 * (3 + 3 + 3 + 2 + 6 + 1),
    # tangential Eface (6) + U_con (3) + B (3) + node electron pressure (1) + U rho (2) + J (3) + eta_local (3)
    # + E_hc (3) + beta (1) + dr (9) + Area (3) + dl (3) + taper (3) read, E (3) + Ehall (3) written
    "edge_total": 8 * (6 + 3 + 3 + 1 + 2 + 3 + 3 + 3 + 1 + 9 + 3 + 3 + 3 + 6),
    # U_s (10) + RHS_s (10) + E (3) + B (3) read, U_s (10) + B (3) + RHS_B (3) written
    "update_all": 8 * (10 + 10 + 3 + 3 + 10 + 3 + 3),
    # B (3) + Badd (3) + W_b (18) + beta (3) + alpha (3) read, B_cell + Bind (6) + J (3) written
    "bcell_jedge": 8 * (3 + 3 + 18 + 3 + 3 + 6 + 3),
    # U_s (10) + J_cell (3) + B_cell (3) + JD (9) + Jac + Area (3, the dt candidate) read,
    # PV_s (10) + U+ (3) + E_hc (3) + beta + U_con (3) written
    "cell_end": 8 * (10 + 3 + 3 + 9 + 1 + 3 + 10 + 3 + 3 + 1 + 3),
    # U_s (2x5) + PV_s u,v,w,p (2x4) + JD (9) + Jac (1) read, RHS_s (2x5) written
    "flux": 8 * (10 + 8 + 9 + 1 + 10),
    # U rho (2) + PV (2x5) + U+ (3) + B_cell (3) + J_cell (3) + Photo + centre (3) + Jac + JD (9) + RHS r/w (2x10)
    "sources": 8 * (2 + 10 + 3 + 3 + 3 + 1 + 3 + 1 + 9 + 20),
    # B (3) + Badd (3) + U+ (3) + JD (9) + Jac (1) read, Eface (9) written
    "eface": 8 * (3 + 3 + 3 + 9 + 1 + 9),
    # Eface (9) + U+ (3) + JD (9) + Jac (1) + B (3) + PV p (2) + U rho (2) + J (3) + eta_local (3) read, E (3) written
    "edge_emf": 8 * (9 + 3 + 9 + 1 + 3 + 2 + 2 + 3 + 3 + 3),
    # U rho (2) + J_cell (3) + B_cell (3) read, Ehc (3) + beta (1) written
    "hall_cell": 8 * (2 + 3 + 3 + 3 + 1),
    # Ehc (3) + beta (1) + dr (9) + Area (3) + dl (3) + B (3) + taper (3) read, Ehall (3) written
    "hall_edge": 8 * (3 + 1 + 9 + 3 + 3 + 3 + 3 + 3),
    # U_s (10) read, PV_s (10) + U+ (3) written
    "pv_uplus": 8 * (10 + 10 + 3),
    # B (3) + Badd (3) + W_b (18) read, B_cell + Bind (6) written
    "bcell": 8 * (3 + 3 + 18 + 6),
    # B (3) + beta (3) + alpha (3) read, J (3) written
    "jedge": 8 * (3 + 3 + 3 + 3),
    # J (3) + W_j (36) read, J_cell (3) written
    "jcell": 8 * (3 + 36 + 3),
    "jcell_t": 8 * (3 + 36 + 3),
    # U_s (10) + B_cell (3) + JD (9) + Jac (1)
    "timestep": 8 * (10 + 3 + 9 + 1),
}
# whole-stage compulsory bytes per cell of this configuration (N_s = 2, Hall, Badd):
# 8 * [2(5N_s+3) + 10 + 60 + 2 + 3 + 18] = 952
STAGE_BYTES = 8 * (2 * (5 * 2 + 3) + 10 + 60 + 2 + 3 + 18)


def measured_traffic(kernel: str, shape=None):
    """HBM bytes per launch of a fused pass from the committed rocprofv3 counter
    summary (2*FETCH_SIZE + WRITE_SIZE, see profiles/shell6_hbm_traffic.md), or None.
    The flux pass is keyed by its launch shape (kind, length)."""
    path = os.path.join(ROOT, "profiles", "shell6_hbm_traffic.md")
    try:
        text = open(path).read()
        blob = text.split("```json", 1)[1].split("```", 1)[0]
        table = json.loads(blob)
    except (OSError, IndexError, ValueError):
        return None
    if kernel == "rhs_cell" and shape is not None:
        return table.get(f"rhs_cell {('column', 'tiled')[shape[0]]} {shape[1]}")
    return table.get(kernel)


def make_case(name: str, resist=None):
    if name == "shell6":
        return cases.make_shell(steps=200, resist=resist)
    if name.startswith("shell6x"):
        return cases.make_shell(scale=float(name[7:]), steps=200, resist=resist)
    if name == "cube48":
        return cases.make_periodic_cube(n=(48, 48, 48), steps=100)
    raise SystemExit(f"unknown workload {name}")


def rel_l2(a, b):
    d = np.sqrt(np.sum((a - b) ** 2))
    n = np.sqrt(np.sum(b ** 2))
    return float(d / n) if n > 0 else float(d)


def cpu_baseline(case, sample_steps: int):
    """Time the CPU restatement on `sample_steps` full steps; its per-point loops
    run on OpenMP threads (OMP_NUM_THREADS, bit-identical to one thread).
    Returns the oracle, the elapsed seconds and the thread count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # noqa: E402
    oracle.build()
    o = oracle.Oracle(case)
    o.prepare()
    t0 = time.perf_counter()
    o.step(sample_steps, case.integrator)
    dt = time.perf_counter() - t0
    return o, dt, int(oracle.lib().orc_threads())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="shell6")
    ap.add_argument("--cpu-sample-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-ops", action="store_true", help="print per-op timings of the fused forms to stderr")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run three device-to-device copies of MIB MiB (a kernel of known traffic for HBM counters)")
    ap.add_argument("--resist-steps", type=int, default=20,
                    help="steps of the shell with the implicit resistive correction (0: skip that measurement)")
    ap.add_argument("--tecplot", type=int, default=1, help="time one Tecplot output of the stepped state (0: skip)")
    ap.add_argument("--rshell-steps", type=int, default=20,
                    help="steps of the resistive shell (4 blocks glued by interfaces, exponential eta) per form; 0 skips it")
    ap.add_argument("--rshell-scale", type=float, default=1.0, help="size factor of the resistive shell (1: 18.9 M cells)")
    ap.add_argument("--ensemble-runs", type=int, default=64,
                    help="members of the batched-ensemble measurement (0: skip; the full configuration has 64)")
    ap.add_argument("--mixed-horizon", type=int, default=200,
                    help="steps of the mixed form against the FP64 form with div B after every step (0: skip)")
    ap.add_argument("--c5-steps", type=int, default=500,
                    help="steps of the three-species resistive shell in the mixed form (0: skip)")
    ap.add_argument("--cut-shell", type=int, default=1,
                    help="also time the six sectors cut into 24 interface-glued blocks (0: skip)")
    ap.add_argument("--patched", type=int, default=1,
                    help="also time the reference's 8-block sample geometry as patch connections against its one-block cube (0: skip)")
    ap.add_argument("--cubed-sphere", type=int, default=64,
                    help="cells per panel side of the cubed-sphere line (6 panels of n^3 cells glued by rotated "
                         "patches, singular corner edges; 0: skip)")
    ap.add_argument("--forms", default="fp64_plain,mixed,fp64_fused",
                    help="comma list of forms to run (fp64_fused is always run: it gives the value)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); every block of the "
              "workload lives on one MI355X and there is no multi-GPU path.", file=sys.stderr)
        sys.exit(2)

    from mpcns_mercury_amd.hip import Mercury  # noqa: E402

    if args.calibrate_copy > 0:
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copies: 3 x {args.calibrate_copy} MiB read + written", file=sys.stderr, flush=True)
        del src, dst

    case = make_case(args.workload)
    ncells = case.ncells()
    print(f"[bench] workload={case.name} blocks={[b.dims for b in case.blocks]} cells={ncells}", file=sys.stderr, flush=True)

    # CPU oracle sample: the baseline timing and the reference state for the error of each form
    cpu, oref = None, None
    if not args.no_cpu:
        oref, tcpu, nthr = cpu_baseline(case, args.cpu_sample_steps)
        cpu = {"value": ncells * args.cpu_sample_steps / tcpu, "unit": "cell-updates/s", "cores": nthr, "kind": "port",
               "sample": f"{args.cpu_sample_steps} full steps of {case.name} ({ncells} cells), CPU oracle, "
                         f"{nthr} OpenMP threads, {tcpu:.1f} s"}
        print(f"[bench] cpu sample {tcpu:.1f}s -> {cpu['value']:.3e} cell-updates/s", file=sys.stderr, flush=True)

    forms = {}
    op_times = {}
    g = None
    fused_ops = ["rhs_cell", "eface_u", "edge_total", "update_all", "bcell_jedge", "jcell_t", "cell_end"]
    wanted = set(args.forms.split(",")) | {"fp64_fused"}
    for form, label in ((0, "fp64_plain"), (2, "mixed"), (1, "fp64_fused")):
        if label not in wanted:
            continue
        if g is not None:
            g.close()
        t0 = time.perf_counter()
        g = Mercury(case)
        g.prepare()
        g.set_form(form)
        info = {"march": g.autotune() if form >= 1 else None}
        if form >= 1:
            kind, length = g.flux_shape()
            info["flux_shape"] = {"kernel": ("k_rhs_cell", "k_rhs_tile")[kind], "kind": kind, "length": length}
        if oref is not None:
            g.step(args.cpu_sample_steps, case.integrator)
            err, emax = {}, {}
            for name in ("U_H", "U_Na", "B_xi", "B_eta", "B_zeta"):
                num, den, mx, scale = 0.0, 0.0, 0.0, 0.0
                for ib in range(case.nb):
                    a_, b_ = g.get(name, ib), oref.get(name, ib)
                    num += float(np.sum((a_ - b_) ** 2))
                    den += float(np.sum(b_ ** 2))
                    mx = max(mx, float(np.abs(a_ - b_).max()))
                    scale = max(scale, float(np.abs(b_).max()))
                err[name] = (num / den) ** 0.5 if den > 0 else num ** 0.5
                emax[name] = mx / scale if scale > 0 else mx
            info["rel_l2_vs_fp64_oracle"] = err
            info["rel_max_vs_fp64_oracle"] = emax
        g.graph_capture(case.integrator)
        g.graph_launch(args.warmup)
        g.synchronize()
        t1 = time.perf_counter()
        g.graph_launch(args.steps)
        g.synchronize()
        el = time.perf_counter() - t1
        dt_last = g.timestep()
        if not (np.isfinite(dt_last) and dt_last > 1e-7):
            raise SystemExit(f"bench: dt underflow after the timed steps (form {label}, dt={dt_last})")
        info["ms_per_step"] = el * 1e3 / args.steps
        info["cell_updates_per_s"] = ncells * args.steps / el
        info["divB_normalised"] = g.divb()[0]
        forms[label] = info
        print(f"[bench] {label}: {info['ms_per_step']:.3f} ms/step {info['cell_updates_per_s']:.4e} cell-updates/s "
              f"divB={info['divB_normalised']:.2e} err={info.get('rel_l2_vs_fp64_oracle')} (setup+run {time.perf_counter() - t0:.1f}s)",
              file=sys.stderr, flush=True)
        if form >= 1:
            # per-kernel device times of the fused passes (HIP events on the solver stream)
            op_times[label] = {op: g.time_op(op, 20) for op in fused_ops}
            if args.profile_ops:
                for op in fused_ops:
                    t = op_times[label][op]
                    print(f"[bench] {label} op {op:12s} {t * 1e3:9.1f} us  {ALGO_BYTES[op] * ncells / (t * 1e-3) / 1e9:8.1f} GB/s",
                          file=sys.stderr)
    # the same six sectors cut into 24 blocks glued by interfaces: the step with
    # inter-block ghost copies in every sync (the uncut workload has none)
    cut = None
    if args.cut_shell and args.workload == "shell6":
        t0 = time.perf_counter()
        cc = cases.make_cut_shell()
        cg = Mercury(cc)
        cg.prepare()
        cut = {"blocks": cc.nb, "cells": cc.ncells(), "split": list(cc.meta["split"])}
        for form, label in ((1, "fp64_fused"), (2, "mixed")):
            cg.set_form(form)
            cg.autotune()
            cg.graph_capture(0)
            cg.graph_launch(args.warmup)
            cg.synchronize()
            t1 = time.perf_counter()
            cg.graph_launch(args.steps)
            cg.synchronize()
            el = time.perf_counter() - t1
            cut[label] = {"ms_per_step": el * 1e3 / args.steps, "cell_updates_per_s": cc.ncells() * args.steps / el,
                          "uncut_ms_per_step": forms[label]["ms_per_step"] if label in forms else None,
                          "divB_normalised": cg.divb()[0]}
            print(f"[bench] cut shell {label}: {cc.nb} blocks, {el * 1e3 / args.steps:.3f} ms/step "
                  f"(uncut {cut[label]['uncut_ms_per_step']}), div B {cut[label]['divB_normalised']:.2e}", file=sys.stderr, flush=True)
        cut["sync_ms"] = {op: cg.time_op(op, 20) for op in ("sync_ucell", "sync_bface", "sync_eedge", "sync_ehall", "sync_jedge",
                                                             "sync_bcell", "sync_jcell")}
        cg.close()
        print(f"[bench] cut shell setup+run {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    # the reference's 8-block sample geometry (Cartesian blocks of one periodic 64^3 cube
    # with every axis permutation and reversal, split faces) as patch connections, under
    # the node-index rule (the reference's halo) and the position rule, against the
    # one-block cube of the same cells
    patched = None
    if args.patched:
        import tarfile
        import tempfile
        t0 = time.perf_counter()
        patched = {}
        with tempfile.TemporaryDirectory() as td:
            with tarfile.open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                           "z0_null_case_geometry.tar.gz")) as tf:
                tf.extractall(td, members=[m for m in tf.getmembers() if "Zone.Identifier" not in m.name])
            from mpcns_mercury_amd.hip import CaseGeometry  # noqa: E402
            geo = CaseGeometry(os.path.join(td, "geometry_3"))
            cube = cases.cartesian_periodic_cube(64, g=4)
            runs = [("one_block_cube", cube)] + [(f"patched_rule{r}", cases.cube_on_geometry(geo, cube, rule=r)) for r in (0, 1)]
            geo.close()
        for label, pc in runs:
            pg = Mercury(pc)
            pg.set_form(1)
            pg.prepare()
            pg.graph_capture(0)
            pg.graph_launch(args.warmup)
            pg.synchronize()
            t1 = time.perf_counter()
            pg.graph_launch(args.steps)
            pg.synchronize()
            el = time.perf_counter() - t1
            patched[label] = {"blocks": pc.nb, "cells": pc.ncells(), "ms_per_step": el * 1e3 / args.steps,
                              "cell_updates_per_s": pc.ncells() * args.steps / el, "divB_normalised": pg.divb()[0]}
            if pc.patches is not None:
                patched[label]["patches"] = int(len(pc.patches))
                patched[label]["sync_ms"] = {op: pg.time_op(op, 20) for op in ("sync_ucell", "sync_bface", "sync_eedge")}
            print(f"[bench] {label}: {pc.nb} blocks, {el * 1e3 / args.steps:.3f} ms/step, div B {patched[label]['divB_normalised']:.2e}",
                  file=sys.stderr, flush=True)
            pg.close()
        print(f"[bench] patched sample setup+run {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    # a whole cubed sphere: six gnomonic panels of n^3 cells (radial x two angular axes)
    # joined by rotated patches, whose eight corner lines are singular edges (three
    # panels meeting), under both staggered-DOF rules, FP64 fused form from a graph
    sphere = None
    if args.cubed_sphere:
        from mpcns_mercury_amd.hip import singular_edges  # noqa: E402
        t0 = time.perf_counter()
        sphere = {}
        n = args.cubed_sphere
        for rule in (0, 1):
            sc = cases.make_cubed_sphere(n, n, rule=rule)
            nsing = len(singular_edges(sc))
            pg = Mercury(sc)
            pg.set_form(1)
            pg.prepare()
            pg.graph_capture(0)
            pg.graph_launch(args.warmup)
            pg.synchronize()
            t1 = time.perf_counter()
            pg.graph_launch(args.steps)
            pg.synchronize()
            el = time.perf_counter() - t1
            label = f"rule{rule}"
            sphere[label] = {"panels": sc.nb, "cells": sc.ncells(), "patches": int(len(sc.patches)), "singular_edges": nsing,
                             "ms_per_step": el * 1e3 / args.steps, "cell_updates_per_s": sc.ncells() * args.steps / el,
                             "divB_normalised": pg.divb()[0],
                             "op_ms": {op: pg.time_op(op, 20) for op in ("singular_emf", "singular_j", "sync_eedge")}}
            print(f"[bench] cubed sphere {n}^3 x 6 {label}: {nsing} singular edges, {el * 1e3 / args.steps:.3f} ms/step, "
                  f"div B {sphere[label]['divB_normalised']:.2e}", file=sys.stderr, flush=True)
            pg.close()
        print(f"[bench] cubed sphere setup+run {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    # visualisation output of the stepped state (Tecplot, the Mercury variable list at the
    # nodes): node means on the device, file written by the host; outside the timed steps
    output = None
    if args.tecplot and g is not None:
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "flow_field.plt")
            t0 = time.perf_counter()
            g.write_tecplot(path, time=0.0)
            dt_out = time.perf_counter() - t0
            output = {"tecplot_s": dt_out, "tecplot_bytes": os.path.getsize(path)}
        print(f"[bench] tecplot output: {dt_out:.3f} s, {output['tecplot_bytes'] / 1e6:.1f} MB", file=sys.stderr, flush=True)
    # the mixed form over the full horizon of the configuration against the FP64 form
    # (bit-identical to the CPU oracle: checked above after the sample steps), with the
    # normalised div B after every step
    horizon = None
    if args.mixed_horizon > 0:
        g.close()
        g = None
        t0 = time.perf_counter()
        gf, gm = Mercury(case), Mercury(case)
        gf.prepare()
        gm.prepare()
        gf.set_form(1)
        gm.set_form(2)
        gf.graph_capture(case.integrator)
        gf.graph_launch(args.mixed_horizon)
        # the mixed run from its step graph; the div B monitor of the B-cell
        # pass gives the maximum of the normalised div B over every step
        gm.set_divb_monitor(True)
        gm.graph_capture(case.integrator)
        gm.divb_monitor(reset=True)
        gm.graph_launch(args.mixed_horizon)
        gm.synchronize()
        _, divb_max, _ = gm.divb_monitor()
        gf.synchronize()
        err, emax = {}, {}
        for name in ("U_H", "U_Na", "B_xi", "B_eta", "B_zeta"):
            num = den = mx = scale = 0.0
            for ib in range(case.nb):
                a_, b_ = gm.get(name, ib), gf.get(name, ib)
                num += float(np.sum((a_ - b_) ** 2))
                den += float(np.sum(b_ ** 2))
                mx = max(mx, float(np.abs(a_ - b_).max()))
                scale = max(scale, float(np.abs(b_).max()))
            err[name] = (num / den) ** 0.5 if den > 0 else num ** 0.5
            emax[name] = mx / scale if scale > 0 else mx
        horizon = {"steps": args.mixed_horizon, "rel_l2_vs_fp64": err, "rel_max_vs_fp64": emax,
                   "divB_normalised_max_over_steps": divb_max, "divB_per_step_us": gm.time_op("divb", 20) * 1e3,
                   "divB_monitor_us": (gm.time_op("bcell_jedge_monitor", 20) - gm.time_op("bcell_jedge", 20)) * 1e3,
                   "reference": "FP64 fused GPU form, bit-identical to the CPU oracle"}
        print(f"[bench] mixed over {args.mixed_horizon} steps: rel-L2 {err}, max div B {divb_max:.2e} "
              f"({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
        gm.close()
        gf.close()
        g = Mercury(case)
        g.prepare()
        g.set_form(1)
        g.graph_capture(case.integrator)
    ms_step = forms["fp64_fused"]["ms_per_step"]
    value = forms["fp64_fused"]["cell_updates_per_s"]
    divb_norm = forms["fp64_fused"]["divB_normalised"]
    err = forms["fp64_fused"].get("rel_l2_vs_fp64_oracle")

    # the same workload with the implicit resistive correction (reference sample window
    # r 0.84-1.04, inver_Rem 0.05): host-driven CG, so not graph captured
    resist = None
    if args.resist_steps > 0 and args.workload.startswith("shell6"):
        g.close()
        rcase = make_case(args.workload, resist=0.05)
        g = Mercury(rcase)
        g.prepare()
        g.set_form(1)
        g.step(2, 0)
        g.synchronize()
        its = []
        t1 = time.perf_counter()
        for _ in range(args.resist_steps):
            g.step(1, 0)
            its.append(g.resist_stats()[1])
        g.synchronize()
        el = time.perf_counter() - t1
        n_unk, _, rel = g.resist_stats()
        resist = {"ms_per_step": el * 1e3 / args.resist_steps, "cell_updates_per_s": ncells * args.resist_steps / el,
                  "unknown_edges": n_unk, "cg_iterations_mean": float(np.mean(its)), "last_rel_residual": rel,
                  "form": "fp64_fused", "inver_Rem": 0.05}
        print(f"[bench] resistive: {resist['ms_per_step']:.3f} ms/step, {n_unk} unknown edges, "
              f"{resist['cg_iterations_mean']:.1f} CG iterations/step, rel residual {rel:.1e}", file=sys.stderr, flush=True)
        g.close()
        g = Mercury(case)
        g.prepare()
        g.set_form(1)
        g.graph_capture(case.integrator)
    # batched ensemble of independent single-fluid + Hall runs (seeded sizes in
    # [32,96]^3): every run's blocks in one context, one launch per pass (per-run
    # physics, dt and status); at half scale also against separate member
    # contexts stepped by one forked graph and one after another
    ensemble = None
    if args.ensemble_runs > 0:
        from mpcns_mercury_amd.hip import Ensemble  # noqa: E402
        t0 = time.perf_counter()
        runs = [cases.make_ensemble_run(r) for r in range(args.ensemble_runs)]
        ecells = sum(c.ncells() for c in runs)
        eg = Mercury(cases.concat_runs(runs))
        eg.prepare()
        ensemble = {"runs": args.ensemble_runs, "cells": ecells, "blocks": sum(c.nb for c in runs),
                    "setup_s": time.perf_counter() - t0, "form": "one context, one launch per pass over every run"}
        for form, label in ((1, "fp64_fused"), (2, "mixed")):
            eg.set_form(form)
            eg.graph_capture(0)
            eg.graph_launch(3)
            eg.synchronize()
            t1 = time.perf_counter()
            eg.graph_launch(20)
            eg.synchronize()
            el = time.perf_counter() - t1
            ensemble[label] = {"ms_per_step": el * 1e3 / 20, "cell_updates_per_s": ecells * 20 / el}
            print(f"[bench] ensemble {label}: {args.ensemble_runs} runs, {ecells} cells in one context: {el * 1e3 / 20:.3f} ms/step "
                  f"({ecells * 20 / el:.3e} cell-updates/s)", file=sys.stderr, flush=True)
        ensemble["mixed_over_fp64"] = ensemble["mixed"]["cell_updates_per_s"] / ensemble["fp64_fused"]["cell_updates_per_s"]
        eg.close()
        del runs
        # half scale: one context vs member contexts (forked graph, and member graphs one after another)
        hruns = [cases.make_ensemble_run(r, scale=0.5) for r in range(args.ensemble_runs)]
        hcells = sum(c.ncells() for c in hruns)
        eg = Mercury(cases.concat_runs(hruns))
        eg.prepare()
        members = [Mercury(c) for c in hruns]
        for m in members:
            m.prepare()
        half = {"runs": args.ensemble_runs, "cells": hcells, "form": "fp64_fused"}
        eg.set_form(1)
        eg.graph_capture(0)
        eg.graph_launch(5)
        for m in members:
            m.set_form(1)
            m.graph_capture(0)
            m.graph_launch(5)
        eg.synchronize()
        for m in members:
            m.synchronize()
        # the same five steps: every run of the one context equals its own member context
        base = 0
        same = True
        for c, m in zip(hruns, members):
            for ib in range(c.nb):
                for name in ("U_H", "B_xi"):
                    same = same and np.array_equal(eg.get(name, base + ib), m.get(name, ib))
            base += c.nb
        half["one_context_equals_members"] = bool(same)
        t1 = time.perf_counter()
        eg.graph_launch(20)
        eg.synchronize()
        half["one_context_ms_per_step"] = (time.perf_counter() - t1) * 1e3 / 20
        ens = Ensemble(members, 0)
        ens.launch(2)
        t1 = time.perf_counter()
        ens.launch(20)
        half["forked_graph_ms_per_step"] = (time.perf_counter() - t1) * 1e3 / 20
        ens.close()
        for m in members:
            m.graph_capture(0)
        t1 = time.perf_counter()
        for _ in range(10):
            for m in members:
                m.graph_launch(1)
        for m in members:
            m.synchronize()
        half["serial_member_graphs_ms_per_step"] = (time.perf_counter() - t1) * 1e3 / 10
        half["one_context_over_serial"] = half["serial_member_graphs_ms_per_step"] / half["one_context_ms_per_step"]
        ensemble["half_scale"] = half
        print(f"[bench] ensemble half scale ({hcells} cells): one context {half['one_context_ms_per_step']:.3f} ms/step, forked graph "
              f"{half['forked_graph_ms_per_step']:.3f}, member graphs one after another {half['serial_member_graphs_ms_per_step']:.3f}; "
              f"runs equal their members: {same}", file=sys.stderr, flush=True)
        for m in members:
            m.close()
        eg.close()
        print(f"[bench] ensemble setup+run {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    # resistive shell: 256 x 192 x 384 cells in 2 x 2 blocks glued by interfaces, two-fluid
    # Hall MHD with the implicit resistive correction on every edge (eta = 1e-3 exp(-(r-1)/0.1))
    rshell = None
    if args.rshell_steps > 0:
        t0 = time.perf_counter()
        rc = cases.make_resistive_shell(scale=args.rshell_scale)
        rcells = rc.ncells()
        rg = Mercury(rc)
        rg.prepare()
        rshell = {"cells": rcells, "blocks": [list(b.dims) for b in rc.blocks], "steps": args.rshell_steps,
                  "setup_s": time.perf_counter() - t0}
        for form, label in ((1, "fp64_fused"), (2, "mixed")):
            rg.set_form(form)
            rg.step(2, 0)
            rg.synchronize()
            its = []
            t1 = time.perf_counter()
            for _ in range(args.rshell_steps):
                rg.step(1, 0)
                its.append(rg.resist_stats()[1])
            rg.synchronize()
            el = time.perf_counter() - t1
            n_unk, _, rel = rg.resist_stats()
            rshell[label] = {"ms_per_step": el * 1e3 / args.rshell_steps, "cell_updates_per_s": rcells * args.rshell_steps / el,
                             "unknown_edges": n_unk, "cg_iterations_mean": float(np.mean(its)), "last_rel_residual": rel,
                             "divB_normalised": rg.divb()[0]}
            print(f"[bench] resistive shell {label}: {rcells} cells, {el * 1e3 / args.rshell_steps:.3f} ms/step, "
                  f"{n_unk} unknown edges, {np.mean(its):.1f} CG iterations/step", file=sys.stderr, flush=True)
        rshell["mixed_over_fp64"] = rshell["mixed"]["cell_updates_per_s"] / rshell["fp64_fused"]["cell_updates_per_s"]
        rshell["species"] = int(rc.phys.n_species)
        if args.c5_steps > 0:
            # the configuration's full run in the mixed form (steps already taken count)
            t1 = time.perf_counter()
            rest = args.c5_steps - (args.rshell_steps + 2) * 2
            its = []
            for _ in range(max(rest, 0)):
                rg.step(1, 0)
                its.append(rg.resist_stats()[1])
            rg.synchronize()
            el = time.perf_counter() - t1
            rshell["mixed_full_run"] = {"steps_total": args.c5_steps, "timed_steps": max(rest, 0),
                                        "ms_per_step": el * 1e3 / max(rest, 1), "cg_iterations_mean": float(np.mean(its)) if its else None,
                                        "divB_normalised": rg.divb()[0],
                                        "he_density_min": float(min(rg.get("U_He", ib)[0].min() for ib in range(rc.nb)))}
            print(f"[bench] resistive shell mixed full run: {args.c5_steps} steps, {el * 1e3 / max(rest, 1):.3f} ms/step, "
                  f"div B {rshell['mixed_full_run']['divB_normalised']:.2e}", file=sys.stderr, flush=True)
        rg.close()
        print(f"[bench] resistive shell setup+run {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    times = op_times["fp64_fused"]
    sync_times = {op: g.time_op(op, 20) for op in ("sync_ucell", "sync_bface", "sync_eedge", "sync_ehall", "sync_jedge",
                                                   "sync_bcell", "sync_jcell")}
    step_graph_ms = g.time_op("step_graph", 10)
    dom = max(times, key=times.get)
    units = ncells
    algo = ALGO_BYTES[dom] * units
    achieved = algo / (times[dom] * 1e-3) / 1e9
    if args.profile_ops:
        for op, t in sync_times.items():
            print(f"[bench] {op:12s} {t * 1e3:9.1f} us", file=sys.stderr)
        print(f"[bench] step (graph) {step_graph_ms:.3f} ms", file=sys.stderr)
    step_gbs = STAGE_BYTES * ncells / (step_graph_ms * 1e-3) / 1e9

    out = {
        "metric": "cell-updates/s (mixed vs FP64 GPU vs CPU) + % HBM roofline, err vs FP64",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{case.name}: two-fluid Hall MHD, FP64 fused GPU form, {case.nb} curvilinear blocks "
                               f"{[list(b.dims) for b in case.blocks]}, forward Euler, CFL 0.4",
                   "cells": ncells, "blocks": case.nb, "form": "fp64_fused", "integrator": "euler"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (measured_traffic(dom, (forms["fp64_fused"]["flux_shape"]["kind"],
                                                        forms["fp64_fused"]["flux_shape"]["length"]))
                                 if args.workload == "shell6" else None), "kernel": dom,
                     "algo_bytes_per_unit": ALGO_BYTES[dom], "units_per_launch": units,
                     "avg_launch_ms": times[dom]},
        "step_roofline": {"stage_bytes_per_cell": STAGE_BYTES, "ms_per_step": step_graph_ms, "achieved_GBs": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "forms": forms,
        "mixed_over_fp64_plain": (forms["mixed"]["cell_updates_per_s"] / forms["fp64_plain"]["cell_updates_per_s"]
                                  if "mixed" in forms and "fp64_plain" in forms else None),
        "err_vs_fp64_oracle": err,
        "divB_normalised": divb_norm,
        "kernel_ms": op_times,
        "resistive": resist,
        "ensemble": ensemble,
        "cut_shell": cut,
        "patched_sample": patched,
        "cubed_sphere": sphere,
        "resistive_shell": rshell,
        "mixed_horizon": horizon,
        "output": output,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
