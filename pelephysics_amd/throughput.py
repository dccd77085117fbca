"""Throughput of one integrator on a synthetic workload, state resident on the device.

    python -m pelephysics_amd.throughput --mech heptane_lu_88sk --cells 131072 \\
        --integrator ReactorBDFAdaptiveMixed --calls 3 [--chained [--graph]] [--rtol 1e-10 --atol 1e-16]

Workloads (pelephysics_amd.workloads): LiDryer -> lidryer_h2air, drm19 ->
drm19_grf (cells rounded to a cube), dodecane_lu -> dodecane_uneven,
heptane_lu_88sk -> heptane_ignition.  Each call integrates dt = 1e-5 s with
the 1-D form on device pointers; without --chained every call restarts from
the initial state, with --chained the calls continue from each other (as a
flow solver's successive chemistry steps do).  Prints one JSON line: cells/s
from the kernel time of each call (HIP events around the launch), the
per-cell work counters (means, and the largest and 99th-percentile step
count of a cell) and the matrix workspace per resident cell.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import workloads as W
from .mechgen.__main__ import load
from .reactor import ReactorBase

GENERATORS = {"LiDryer": W.lidryer_h2air, "dodecane_lu": W.dodecane_uneven,
              "heptane_lu_88sk": W.heptane_ignition, "dodecane_lu_qss": W.dodecane_uneven,
              "CH4_lean_qss": W.premixed, "LuEthylene_qss": W.premixed}


def workload(mech: str, ncells: int) -> W.CellState:
    """State for a compiled mechanism name (a "_table" suffix selects the
    table-driven form of the same mechanism)."""
    base = mech[:-len("_table")] if mech.endswith("_table") else mech
    m = load(base)
    if base == "drm19":
        side = max(1, round(ncells ** (1.0 / 3.0)))
        return W.drm19_grf(m, side)
    return GENERATORS[base](m, ncells)


def main(argv=None) -> dict:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mech", default="heptane_lu_88sk")
    ap.add_argument("--cells", type=int, default=131072)
    ap.add_argument("--integrator", default="ReactorBDFAdaptiveMixed")
    ap.add_argument("--calls", type=int, default=3)
    ap.add_argument("--chained", action="store_true")
    ap.add_argument("--graph", action="store_true",
                    help="with --chained: all calls in one pp_chem_react_1d_chained (HIP graph replay)")
    ap.add_argument("--dt", type=float, default=1e-5)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--atol", type=float, default=1e-16)
    ap.add_argument("--waves-per-simd", type=int, default=0)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--lu-batch", type=int, default=0)
    ap.add_argument("--lu-wait", type=int, default=0)
    ap.add_argument("--cells-per-wave", type=int, default=0,
                    help="batched forms: 0 one cell per lane; 8, 4, 2, 1 a group of 64/k lanes per cell")
    a = ap.parse_args(argv)

    import torch
    st = workload(a.mech, a.cells)
    m = load(a.mech[:-len("_table")] if a.mech.endswith("_table") else a.mech)
    dev = torch.device("cuda:0")
    R = ReactorBase.create(a.integrator)
    R.init(2, st.ncells, mechanism=a.mech, rtol=a.rtol, atol=a.atol, batched_waves_per_cu=a.waves_per_cu,
           batched_lu_batch=a.lu_batch, batched_lu_wait=a.lu_wait, batched_waves_per_simd=a.waves_per_simd,
           batched_cells_per_wave=a.cells_per_wave)
    y0 = torch.from_numpy(st.y_1d()).to(dev)
    x0 = torch.from_numpy(st.rE.copy()).to(dev)
    y, x = y0.clone(), x0.clone()
    ys = torch.zeros(st.ncells * m.n_species, dtype=torch.float64, device=dev)
    xs = torch.zeros(st.ncells, dtype=torch.float64, device=dev)
    kms, per_call = [], []
    t0 = time.perf_counter()
    if a.chained and a.graph:
        R.react_1d_chained(y, ys, x, xs, a.dt, a.calls)
        s = R.stats()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out = {"mech": a.mech, "integrator": a.integrator, "cells": st.ncells, "dt": a.dt, "calls": a.calls,
               "chained": True, "graph": True, "kernel_ms_total": s["kernel_ms"], "wall_s": wall,
               "cells_per_s": st.ncells * a.calls / wall,
               "finite": bool(torch.isfinite(y).all().item())}
        R.close()
        print(json.dumps(out))
        return out
    for call in range(a.calls):
        if not a.chained:
            y.copy_(y0)
            x.copy_(x0)
        R.react_1d(y, ys, x, xs, a.dt)
        s = R.stats()
        kms.append(s["kernel_ms"])
        steps = R.cell_steps(st.ncells)
        per_call.append(dict({k: s[k] / max(1, s["ncells"]) for k in ("nsteps", "nrhs", "njac", "nlu")},
                             max_steps=int(steps.max()), p99_steps=float(np.percentile(steps, 99))))
        print(f"[throughput] call {call + 1}/{a.calls}: kernel {s['kernel_ms']:.1f} ms, "
              f"{s['nsteps'] / max(1, s['ncells']):.1f} steps/cell (max {per_call[-1]['max_steps']}), "
              f"failed {s['nfailed']}",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {
        "mech": a.mech, "integrator": a.integrator, "cells": st.ncells, "dt": a.dt,
        "rtol": a.rtol, "atol": a.atol, "calls": a.calls, "chained": a.chained,
        "cells_per_s": st.ncells / (float(np.median(kms)) * 1e-3),
        "cells_per_s_wall": st.ncells * a.calls / wall,
        "kernel_ms": kms, "wall_s": wall, "per_cell": per_call,
        "workspace_bytes_per_resident_cell": R.workspace_rows() * 8,
        "finite": bool(torch.isfinite(y).all().item()),
    }
    R.close()
    print(json.dumps(out))
    return out


if __name__ == "__main__":
    main()
