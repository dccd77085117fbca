"""Benchmark: drm19 chemistry on 64^3 cells, dt = 1e-5 s, one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode mixed|fp64]

A step is one ``react`` call (box/Array4 form, h-reactor) over the whole 64^3
grid of the drm19 Gaussian-random-field workload, state resident in HBM; the
pristine state is restored (device-to-device copy) before every step.  The
headline form is the batched adaptive BDF (persistent work-queue kernel, FP32
Newton-matrix LU); the plain FP64 adaptive BDF (one lane per cell) is timed
beside it.  Prints one JSON line.
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

MECH = "drm19"
GRID = 64
DT = 1e-5
# Tolerances of the adaptive BDF: the config's rtol = 1e-8 with atol = 1e-10 on
# rhoY leaves minor species above the 1e-6 relative error criterion; the
# absolute tolerance is lowered until every cell meets it.
RTOL, ATOL = 1e-10, 1e-16
P_FP64 = 78.6  # TFLOP/s, MI355X FP64 (vector and matrix)
P_FP32 = 157.3  # TFLOP/s, MI355X FP32


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TUNABLES = ("batched_waves_per_cu", "batched_lu_batch", "batched_lu_wait", "batched_box_direct",
            "batched_waves_per_simd")


def run_gpu(ident, st, m, steps, warmup, rtol, atol, mech=MECH, launch=None):
    import torch
    from pelephysics_amd.reactor import ReactorBase
    dev = torch.device("cuda:0")
    ns = m.n_species
    n = GRID
    R = ReactorBase.create(ident)
    R.init(2, st.ncells, mechanism=mech, rtol=rtol, atol=atol, **(launch or {}))

    def soa(a, nc):
        return torch.from_numpy(np.ascontiguousarray(a.T.reshape(nc, n, n, n))).to(dev)

    rY0 = soa(st.rY, ns)
    T0 = soa(st.T[:, None], 1)
    rE0 = soa(st.rE[:, None], 1)
    rY, T, rE = rY0.clone(), T0.clone(), rE0.clone()
    rYs = torch.zeros_like(rY0)
    rEs = torch.zeros_like(rE0)
    FC = torch.zeros_like(T0)

    def step():
        rY.copy_(rY0)
        T.copy_(T0)
        rE.copy_(rE0)
        return R.react(rY, rYs, T, rE, rEs, FC, DT)[0]

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        kms.append(R.stats()["kernel_ms"])
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    s = R.stats()
    out = {
        "elapsed": el, "kernel_ms": float(np.mean(kms)), "stats": s,
        "T": T.reshape(-1).cpu().numpy(), "rY": rY.reshape(ns, -1).cpu().numpy().T,
    }
    R.close()
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_drm19_64cubed.json")


def measured_traffic(ident, kernels, launch):
    """HBM bytes per launch of the integration kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (profiles/), or None when they
    were taken on other kernels or another launch shape."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get(ident)
    if e is None or e.get("kernels") != kernels or e.get("launch", {k: 0 for k in TUNABLES}) != launch:
        return None
    return float(e["bytes_per_launch"])


def copy_only(nbytes):
    """A plain device-to-device copy of nbytes, the calibration kernel of the
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] copied {nbytes} bytes")


def roofline(m, s, kernel_ms, mixed):
    from pelephysics_amd.opcount import counts, work_per_cell
    oc = counts(m)
    w = work_per_cell(oc, s["nrhs"], s["njac"], s["nlu"], s["nnewton"])  # totals over all cells
    w64 = w["rhs"] + w["jac"] + w["solve"] + (0.0 if mixed else w["lu"])
    w32 = w["lu"] if mixed else 0.0
    t = kernel_ms * 1e-3
    achieved = (w64 + w32 * P_FP64 / P_FP32) / t / 1e12
    return {"bound": "mfma", "achieved": achieved, "peak": P_FP64, "unit": "TFLOP/s",
            "frac": achieved / P_FP64, "traffic": None, "algorithmic_bytes": None,
            "flops_per_launch_fp64": w64, "flops_per_launch_fp32": w32,
            "F_rhs": oc.rhs, "F_jac": oc.jac}


def cpu_baseline(m, st):
    """The oracle's adaptive BDF (C restatement, OpenMP over cells) on the whole
    64^3 grid: about 7 s on 16 host threads."""
    import oracle
    O = oracle.Mech(m)
    idx = np.arange(st.ncells)
    from pelephysics_amd.workloads import subsample
    sub = subsample(st, idx)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or 1))
    z = np.zeros(sub.ncells)
    src = np.zeros(sub.ncells * m.n_species)
    p = oracle.adapt_params(rtol=RTOL, atol=ATOL)
    # warm-up on a few cells
    y = sub.y_1d()[: 64 * (m.n_species + 1)].copy()
    O.react_1d("adaptive", y, src[: 64 * m.n_species].copy(), sub.rE[:64].copy(), z[:64].copy(), DT,
               params=p, nthreads=threads)
    y = sub.y_1d()
    rX = sub.rE.copy()
    t0 = time.perf_counter()
    O.react_1d("adaptive", y, src, rX, z, DT, params=p, nthreads=threads)
    el = time.perf_counter() - t0
    return {"value": sub.ncells / el, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"all {sub.ncells} cells of the 64^3 drm19 grid, one call, adaptive BDF "
                      f"rtol={RTOL:g} atol={ATOL:g}, {el:.1f} s"}


def adjoint_measure(m):
    """Adjoint workload: drm19, 4096 fresh CH4/air cells (T ~ U[1100,1500] K,
    phi ~ U[0.6,1.4], 1 atm), 10 ReactorBDF calls of 1e-5 s, d(T, Y_CO, Y_OH)/d ln A_j
    for all reactions.  Time of the adjoint kernel (forward with checkpoints +
    reverse sweep) against the forward pass alone (the same 10 ReactorBDF calls)."""
    import torch
    from pelephysics_amd.reactor import ReactorBase
    from pelephysics_amd import workloads as W
    dev = torch.device("cuda:0")
    st = W.drm19_fresh(m, 4096)
    n = m.n_species + 1
    y0 = torch.from_numpy(st.y_1d().reshape(st.ncells, n)).to(dev)
    rX0 = torch.from_numpy(st.rE.copy()).to(dev)
    R = ReactorBase.create("ReactorBDF")
    R.init(2, st.ncells, mechanism=MECH)
    R.adjoint_1d(y0.clone(), rX0, 10, DT)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out, grad = R.adjoint_1d(y0.clone(), rX0, 10, DT)
    torch.cuda.synchronize()
    adj_s = time.perf_counter() - t0
    adj_kms = R.stats()["kernel_ms"]
    ys = torch.zeros(st.ncells * m.n_species, dtype=torch.float64, device=dev)
    zs = torch.zeros(st.ncells, dtype=torch.float64, device=dev)
    y, rX = y0.clone().reshape(-1), rX0.clone()
    R.react_1d(y, ys, rX, zs, DT)
    torch.cuda.synchronize()
    y, rX = y0.clone().reshape(-1), rX0.clone()
    t0 = time.perf_counter()
    fwd_kms = 0.0
    for _ in range(10):
        R.react_1d(y, ys, rX, zs, DT)
        fwd_kms += R.stats()["kernel_ms"]
    torch.cuda.synchronize()
    fwd_s = time.perf_counter() - t0
    R.close()
    return {"workload": "drm19 4096 fresh cells, 10 ReactorBDF calls of 1e-5 s, d(T,Y_CO,Y_OH)/d lnA_j, 84 reactions",
            "adjoint_kernel_ms": adj_kms, "adjoint_wall_ms": adj_s * 1e3,
            "forward_kernel_ms": fwd_kms, "forward_wall_ms": fwd_s * 1e3,
            "time_ratio": adj_kms / fwd_kms if fwd_kms > 0 else None,
            "checkpoint_bytes": st.ncells * 10 * n * 8, "state_bytes": st.ncells * n * 8,
            "grad_bytes": st.ncells * m.n_reactions * 3 * 8,
            "finite": bool(torch.isfinite(grad).all().item())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["mixed", "batched-fp64", "fp64"], default="mixed")
    ap.add_argument("--impl", choices=["tuned", "unrolled", "table"], default="tuned",
                    help="drm19 kernels of the measured form: the autotuner's choice (pelephysics_amd/tuning), "
                         "generated straight-line code or constant tables")
    ap.add_argument("--plain-impl", choices=["tuned", "unrolled", "table"], default="tuned",
                    help="drm19 kernels of the plain FP64 form")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only copy this many bytes on the device (counter calibration)")
    a = ap.parse_args()
    if a.copy_bytes:
        copy_only(a.copy_bytes)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported "
              "(cells are independent: shard the grid across processes).", file=sys.stderr)
        sys.exit(2)
    from pelephysics_amd.mechgen.__main__ import load
    from pelephysics_amd import workloads as W
    m = load(MECH)
    st = W.drm19_grf(m, GRID)
    ncells = st.ncells
    mixed = a.mode == "mixed"
    ident = {"mixed": "ReactorBDFAdaptiveMixed", "batched-fp64": "ReactorBDFBatchedFP64",
             "fp64": "ReactorBDFAdaptive"}[a.mode]
    log(f"[bench] {ident} drm19 {GRID}^3 = {ncells} cells, warmup {a.warmup}, steps {a.steps}")
    def choose(kind, integrator):
        """(kernels, launch parameters, tuning entry or None) of one measured form."""
        if kind == "tuned":
            from pelephysics_amd.autotune import lookup
            e = lookup(MECH, integrator, "box")
            if e:
                return (e["kernels"], {k: e.get(k, 0) for k in TUNABLES},
                        {k: e[k] for k in e if k != "candidates"})
            return MECH, {k: 0 for k in TUNABLES}, None
        return {"unrolled": MECH, "table": MECH + "_table"}[kind], {k: 0 for k in TUNABLES}, None

    kern, launch, tuning = choose(a.impl, ident)
    pkern, plaunch, ptuning = choose(a.plain_impl, "ReactorBDFAdaptive")
    head = run_gpu(ident, st, m, a.steps, a.warmup, RTOL, ATOL, kern, launch)
    value = ncells * a.steps / head["elapsed"]
    log(f"[bench] {ident}: {value:.4g} cells/s, kernel {head['kernel_ms']:.2f} ms")
    extra = {}
    if a.mode != "fp64" and not a.no_plain:
        plain = run_gpu("ReactorBDFAdaptive", st, m, max(1, a.steps // 2), 1, RTOL, ATOL, pkern, plaunch)
        pv = ncells * max(1, a.steps // 2) / plain["elapsed"]
        Yp = plain["rY"] / plain["rY"].sum(1, keepdims=True)
        Ym = head["rY"] / head["rY"].sum(1, keepdims=True)
        eT = np.abs(head["T"] - plain["T"]) / (1e-6 * np.abs(plain["T"]) + 1e-10)
        eY = np.abs(Ym - Yp) / (1e-6 * np.abs(Yp) + 1e-10)
        extra = {"plain_fp64_cells_per_s": pv, "speedup_vs_plain_fp64": value / pv,
                 "max_err_ratio_vs_plain_fp64": float(max(eT.max(), eY.max()))}
        log(f"[bench] plain FP64: {pv:.4g} cells/s; speedup {value / pv:.3f}")
    s = head["stats"]
    rl = roofline(m, s, head["kernel_ms"], mixed)
    rl["traffic"] = measured_traffic(ident, kern, launch)
    rl["algorithmic_bytes"] = (3 * m.n_species + 6) * 8 * ncells  # (3N+6)*8 B per cell: read+write once
    adj = None if a.no_adjoint else adjoint_measure(m)
    if adj:
        log(f"[bench] adjoint: {adj['adjoint_kernel_ms']:.1f} ms vs forward {adj['forward_kernel_ms']:.1f} ms")
    cpu = None if a.no_cpu else cpu_baseline(m, st)
    if cpu:
        log(f"[bench] cpu baseline {cpu['value']:.4g} cells/s on {cpu['cores']} threads")
    ms = head["elapsed"] / a.steps * 1e3
    line = {
        "metric": "cells/s drm19 64^3 dt=1e-5s, batched vs plain GPU BDF at <=1e-6 rel err in T,Y",
        "value": value, "unit": "cells/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32" if mixed else "f64", "data": "synthetic",
        "config": {"workload": f"drm19 {GRID}^3 Gaussian-random-field mixture fraction x progress "
                               f"variable, h-reactor, dt=1e-5 s, adaptive BDF rtol={RTOL:g} "
                               f"atol={ATOL:g}",
                   "integrator": ident, "kernels": kern, "launch": launch, "tuning": tuning,
                   "plain_kernels": pkern, "plain_tuning": ptuning,
                   "cells": ncells, "parallelism": "single GPU"},
        "per_cell": {k: s[k] / ncells for k in ("nsteps", "nrhs", "njac", "nlu", "nnewton", "nreject")},
        "kernel_ms": head["kernel_ms"],
        "roofline": rl, "cpu_baseline": cpu,
    }
    line["adjoint"] = adj
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
