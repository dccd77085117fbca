"""Benchmark: drm19 chemistry on 64^3 cells, dt = 1e-5 s, one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode mixed|fp64]

A step is one ``react`` call (box/Array4 form, h-reactor, dt = 1e-5 s) over
the whole 64^3 grid of the drm19 Gaussian-random-field workload, state
resident in HBM; the steps continue from each other on the same buffers (the
state evolves as under a flow solver).  The headline form is the batched
adaptive BDF (persistent work-queue kernel, FP32 Newton-matrix LU); the plain
FP64 adaptive BDF (one lane per cell) runs the same sequence of calls beside
it.  Also reported: the first call on a fresh handle (no cost history), the
rate with the pristine state restored before every call (the queue order of an
identical call: a best case), the error of the first call against the oracle
at rtol 1e-13 on every cell, and the oracle's CPU rate.  Prints one JSON line.
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
            "batched_waves_per_simd", "batched_cells_per_wave")


def run_gpu(ident, st, m, steps, warmup, rtol, atol, mech=MECH, launch=None, replay_steps=0):
    """Times one integrator on the 64^3 grid (box form, device arrays):

    * call 1 from the pristine state on a fresh handle (no cost history: the
      work queue visits the cells in index order) -- its result is the one
      checked against the oracle;
    * warmup - 1 more calls, then `steps` timed calls, each continuing from the
      state the previous call left on the same buffers (the state evolves as
      in a flow solver's chemistry steps; the queue takes the cost order of
      the previous call);
    * with replay_steps, that many calls each restoring the pristine state
      first (the queue order then comes from an identical call: a best case).
    """
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
    work = {k: 0 for k in ("nsteps", "nrhs", "njac", "nlu", "nnewton", "nreject")}

    def call(tally=None):
        R.react(rY, rYs, T, rE, rEs, FC, DT)
        s = R.stats()
        if tally is not None:
            for k in work:
                tally[k] += s[k]
            tally["kernel_ms"] += s["kernel_ms"]
            tally["max_steps_cell"] = max(tally["max_steps_cell"], s["max_steps_cell"])
        return s

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    first_stats = call()
    torch.cuda.synchronize()
    first_s = time.perf_counter() - t0
    out = {"first_ms": first_s * 1e3, "first_kernel_ms": first_stats["kernel_ms"], "first_stats": first_stats,
           "T": T.reshape(-1).cpu().numpy(), "rY": rY.reshape(ns, -1).cpu().numpy().T}
    for _ in range(max(0, warmup - 1)):
        call()
    torch.cuda.synchronize()
    tally = dict(work, kernel_ms=0.0, max_steps_cell=0)
    t0 = time.perf_counter()
    for _ in range(steps):
        call(tally)
    torch.cuda.synchronize()
    out["elapsed"] = time.perf_counter() - t0
    out["tally"] = tally
    out["kernel_ms"] = tally["kernel_ms"] / max(steps, 1)
    if replay_steps:
        rY.copy_(rY0), T.copy_(T0), rE.copy_(rE0)
        call()  # sets the cost history of the pristine state
        rt = dict(work, kernel_ms=0.0, max_steps_cell=0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(replay_steps):
            rY.copy_(rY0), T.copy_(T0), rE.copy_(rE0)
            call(rt)
        torch.cuda.synchronize()
        out["replay_ms"] = (time.perf_counter() - t0) / replay_steps * 1e3
        out["replay_kernel_ms"] = rt["kernel_ms"] / replay_steps
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
    """FP64-equivalent flops of the integration kernels over kernel_ms, from
    the per-cell work counters s (totals over the timed calls) and the static
    op counts of the generator.  The work is FP64 vector arithmetic (VALU); the
    FP64 vector and FP64 matrix peaks of the MI355X are the same 78.6 TFLOP/s,
    so the compute roof is that figure either way."""
    from pelephysics_amd.opcount import counts, work_per_cell
    oc = counts(m)
    w = work_per_cell(oc, s["nrhs"], s["njac"], s["nlu"], s["nnewton"])  # totals over all cells
    w64 = w["rhs"] + w["jac"] + w["solve"] + (0.0 if mixed else w["lu"])
    w32 = w["lu"] if mixed else 0.0
    t = kernel_ms * 1e-3
    achieved = (w64 + w32 * P_FP64 / P_FP32) / t / 1e12
    return {"bound": "mfma", "compute_unit": "FP64 VALU (FP64 vector peak = FP64 matrix peak on MI355X)",
            "achieved": achieved, "peak": P_FP64, "unit": "TFLOP/s",
            "frac": achieved / P_FP64, "traffic": None, "algorithmic_bytes": None,
            "flops_per_launch_fp64": w64, "flops_per_launch_fp32": w32,
            "F_rhs": oc.rhs, "F_jac": oc.jac}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_errors(m, st, outs):
    """The oracle's adaptive BDF at rtol 1e-13 / atol 1e-20 (the FP64 truth
    substitute) on every cell of the grid, one call from the pristine state;
    per output the largest |x - x_ref| / (1e-6 |x_ref| + 1e-10) over cells and
    components (T and Y_k) and the number of cells above 1."""
    import oracle
    O = oracle.Mech(m)
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), os.cpu_count() or 1))
    y = st.y_1d()
    z = np.zeros(st.ncells)
    t0 = time.perf_counter()
    O.react_1d("adaptive", y, np.zeros(st.ncells * m.n_species), st.rE.copy(), z, DT,
               params=oracle.adapt_params(rtol=1e-13, atol=1e-20), nthreads=threads)
    el = time.perf_counter() - t0
    ref = y.reshape(st.ncells, m.n_species + 1)
    rT, rYm = ref[:, -1], ref[:, :-1] / ref[:, :-1].sum(1, keepdims=True)
    res = {"oracle": "adaptive BDF rtol=1e-13 atol=1e-20, all cells", "oracle_s": el}
    for name, o in outs.items():
        Ym = o["rY"] / o["rY"].sum(1, keepdims=True)
        eT = np.abs(o["T"] - rT) / (1e-6 * np.abs(rT) + 1e-10)
        eY = (np.abs(Ym - rYm) / (1e-6 * np.abs(rYm) + 1e-10)).max(1)
        e = np.maximum(eT, eY)
        res[name] = {"max_err_ratio": float(e.max()), "cells_over_criterion": int((e > 1).sum())}
    return res


def cpu_baseline(m, st, runs=5):
    """The oracle's adaptive BDF (C restatement, OpenMP over cells, dynamic
    schedule) on a sample of the grid: every 8th cell (32768 cells), one
    warm-up and the median of `runs` timed calls at the bench tolerances."""
    import oracle
    O = oracle.Mech(m)
    from pelephysics_amd.workloads import subsample
    sub = subsample(st, np.arange(0, st.ncells, 8))
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), os.cpu_count() or 1))
    z = np.zeros(sub.ncells)
    src = np.zeros(sub.ncells * m.n_species)
    p = oracle.adapt_params(rtol=RTOL, atol=ATOL)
    times = []
    for r in range(runs + 1):
        y = sub.y_1d()
        rX = sub.rE.copy()
        t0 = time.perf_counter()
        O.react_1d("adaptive", y, src, rX, z, DT, params=p, nthreads=threads)
        if r > 0:
            times.append(time.perf_counter() - t0)
    el = float(np.median(times))
    return {"value": sub.ncells / el, "unit": "cells/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"every 8th cell of the 64^3 drm19 grid ({sub.ncells} cells), one call each, adaptive BDF "
                      f"rtol={RTOL:g} atol={ATOL:g}; median of {runs} calls after one warm-up: {el:.2f} s"}


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
    head = run_gpu(ident, st, m, a.steps, a.warmup, RTOL, ATOL, kern, launch, replay_steps=a.steps)
    value = ncells * a.steps / head["elapsed"]
    log(f"[bench] {ident}: {value:.4g} cells/s chained, kernel {head['kernel_ms']:.2f} ms; first call "
        f"{head['first_ms']:.1f} ms, replay {head['replay_ms']:.1f} ms")
    extra = {"first_call_ms": head["first_ms"], "first_call_cells_per_s": ncells / head["first_ms"] * 1e3,
             "replay_ms": head["replay_ms"], "replay_cells_per_s": ncells / head["replay_ms"] * 1e3}
    outs = {ident: head}
    if a.mode != "fp64" and not a.no_plain:
        plain = run_gpu("ReactorBDFAdaptive", st, m, a.steps, a.warmup, RTOL, ATOL, pkern, plaunch)
        pv = ncells * a.steps / plain["elapsed"]
        outs["ReactorBDFAdaptive"] = plain
        Yp = plain["rY"] / plain["rY"].sum(1, keepdims=True)
        Ym = head["rY"] / head["rY"].sum(1, keepdims=True)
        eT = np.abs(head["T"] - plain["T"]) / (1e-6 * np.abs(plain["T"]) + 1e-10)
        eY = np.abs(Ym - Yp) / (1e-6 * np.abs(Yp) + 1e-10)
        extra.update({"plain_fp64_cells_per_s": pv, "speedup_vs_plain_fp64": value / pv,
                      "plain_fp64_first_call_ms": plain["first_ms"],
                      "speedup_first_call": plain["first_ms"] / head["first_ms"],
                      "max_err_ratio_vs_plain_fp64": float(max(eT.max(), eY.max()))})
        log(f"[bench] plain FP64: {pv:.4g} cells/s chained; speedup {value / pv:.3f}, first call "
            f"{plain['first_ms'] / head['first_ms']:.3f}")
    s = head["tally"]
    rl = roofline(m, s, head["tally"]["kernel_ms"], mixed)
    rl["traffic"] = measured_traffic(ident, kern, launch)
    rl["algorithmic_bytes"] = (3 * m.n_species + 6) * 8 * ncells  # (3N+6)*8 B per cell: read+write once
    adj = None if a.no_adjoint else adjoint_measure(m)
    if adj:
        log(f"[bench] adjoint: {adj['adjoint_kernel_ms']:.1f} ms vs forward {adj['forward_kernel_ms']:.1f} ms")
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(m, st)
        log(f"[bench] cpu baseline {cpu['value']:.4g} cells/s on {cpu['cores']} threads")
        extra["error_vs_oracle"] = oracle_errors(m, st, outs)
        log(f"[bench] error vs oracle: {extra['error_vs_oracle']}")
    ms = head["elapsed"] / a.steps * 1e3
    line = {
        "metric": "cells/s drm19 64^3 dt=1e-5s, batched vs plain GPU BDF at <=1e-6 rel err in T,Y",
        "value": value, "unit": "cells/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32" if mixed else "f64", "data": "synthetic",
        "config": {"workload": f"drm19 {GRID}^3 Gaussian-random-field mixture fraction x progress "
                               f"variable, h-reactor, dt=1e-5 s per call, calls chained on the same "
                               f"buffers, adaptive BDF rtol={RTOL:g} atol={ATOL:g}",
                   "integrator": ident, "kernels": kern, "launch": launch, "tuning": tuning,
                   "plain_kernels": pkern, "plain_tuning": ptuning,
                   "cells": ncells, "parallelism": "single GPU"},
        "per_cell": {k: s[k] / ncells / a.steps for k in ("nsteps", "nrhs", "njac", "nlu", "nnewton", "nreject")},
        "max_steps_cell": s["max_steps_cell"],
        "kernel_ms": head["kernel_ms"],
        "roofline": rl, "cpu_baseline": cpu,
    }
    line["adjoint"] = adj
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
