"""Autotuner for the kernel implementation and launch shape of an integrator.

    python -m pelephysics_amd.autotune --mech drm19 --integrator ReactorBDFAdaptiveMixed \\
        --form box --cells 262144 [--out tuning.json]

For one (mechanism, integrator, form) it times every candidate on a synthetic
state of the mechanism's workload (pelephysics_amd.throughput.workload) and
keeps the fastest:

* kernels: the compiled implementations of the mechanism, the generated
  straight-line code ("<mech>") and the constant-memory tables
  ("<mech>_table"), where both exist;
* batched_waves_per_simd / batched_waves_per_cu (batched forms only): the
  kernel build held to 2 waves per SIMD (8, 6 or 4 persistent waves per
  compute unit) or to 1 (4 or 3 per CU).  Fewer waves contend less for the
  caches that back private memory but hide less latency; the 1-wave build
  spills less;
* batched_lu_batch / batched_lu_wait (batched forms only): how many waiting
  lanes, or how many trips of waiting, make a wave factor its lanes' Newton
  matrices together (16 lanes or 4 trips; 64 lanes or 8 trips; 64 lanes or
  32 trips, i.e. nearly always the whole wave at once);
* batched_box_direct (batched forms, box calls): the box kernel itself (1) or
  box -> flat state, the 1-D kernel, flat -> box (0);
* batched_cells_per_wave (batched forms, table-driven kernels): one cell per
  lane (0) or a group of 8, 16, 32 or 64 lanes per cell (8, 4, 2, 1 cells per
  wave; the sizes a mechanism's kernels are built for), with 8 or 4 waves
  per CU and the LU service serving each group at once (1 cell, 1 trip) or
  all of a wave's waiting groups together (64, 8 or 32 trips).

Each candidate runs `calls` calls on the same cells and arrays; its time is
the kernel time of the last one (the queue in the cost order of the call
before).  By default every call restarts from the same initial state
(`--protocol replay`); with `--protocol chained` each call continues from the
state the previous one left, as a flow solver's chemistry steps do (the
benchmark's protocol), and the candidates are compared on the last call of
that chain.  A candidate is eligible only if every cell integrates and its
result agrees with the first candidate's to |x - x_ref| <= 1e-6 |x_ref| +
1e-10 on mass fractions and T (the launch shapes do not change any cell's
arithmetic, so a disagreement means a faulty build).  Results go to a JSON table
keyed by the GPU architecture (gcnArchName) and "<mech>|<integrator>|<form>";
``ReactorBase.init(..., tuned="box" or "1d")`` and bench.py read it back
(pelephysics_amd/tuning/<arch>.json unless PELE_CHEM_TUNING names another
file).
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# mechanisms compiled from their tables only, under their plain names
TABLE_ONLY = ("dodecane_lu", "heptane_lu_88sk")
BATCHED = ("ReactorBDFAdaptiveMixed", "ReactorBDFBatchedFP64")
OCCUPANCY = ((2, 8), (2, 6), (2, 4), (1, 4), (1, 3))  # (waves per SIMD, waves per CU)
PLAIN_OCCUPANCY = (0, 2, 4)  # plain adaptive BDF builds: compiler choice, 2 or 4 waves per SIMD
LU_SERVICE = ((16, 4), (64, 8), (64, 32))
GROUPS = (8, 4, 2, 1)  # cells per wave of the group kernel
GROUP_WAVES = (0, 4)  # persistent waves per CU of the group kernel (0: every resident wave)
GROUP_LU = ((1, 1), (64, 8), (64, 32))


def arch() -> str:
    import torch
    return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]


def table_path(a: str | None = None) -> str:
    return os.environ.get("PELE_CHEM_TUNING") or os.path.join(HERE, "tuning", f"{a or 'gfx950'}.json")


def key(mech: str, integrator: str, form: str) -> str:
    return f"{mech}|{integrator}|{form}"


def lookup(mech: str, integrator: str, form: str = "box", a: str | None = None) -> dict | None:
    """The tuned entry for (mechanism, integrator, form), or None."""
    p = table_path(a)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get(key(mech, integrator, form))


def candidates(mech: str, integrator: str, form: str = "1d") -> list:
    from .reactor import mechanisms
    have = set(mechanisms())
    impls = [k for k in (mech, mech + "_table") if k in have]
    if integrator not in BATCHED:
        # the plain adaptive BDF is also built held to 2 and 4 waves per SIMD
        # for mechanisms whose cells fit in registers (NEQ <= 32: no workspace)
        from .mechgen.__main__ import load
        fits = load(mech).n_species + 1 <= 32
        occs = PLAIN_OCCUPANCY if integrator == "ReactorBDFAdaptive" and fits else (0,)
        return [{"kernels": k, "batched_waves_per_simd": o, "batched_waves_per_cu": 0, "batched_lu_batch": 0,
                 "batched_lu_wait": 0, "batched_box_direct": 0} for k in impls for o in occs]
    routes = (0, 1) if form == "box" else (0,)
    lane = [{"kernels": k, "batched_waves_per_simd": o, "batched_waves_per_cu": w, "batched_lu_batch": b,
             "batched_lu_wait": t, "batched_box_direct": r, "batched_cells_per_wave": 0}
            for k in impls for o, w in OCCUPANCY for b, t in LU_SERVICE for r in routes
            if not (o == 1 and r == 1)]  # the box kernel is built at 2 waves per SIMD only
    # group kernels: table-driven implementations only (a size a mechanism is
    # not built for is rejected at init and skipped)
    tables = [k for k in impls if k.endswith("_table") or k in TABLE_ONLY]
    group = [{"kernels": k, "batched_waves_per_simd": 0, "batched_waves_per_cu": w, "batched_lu_batch": b,
              "batched_lu_wait": t, "batched_box_direct": 0, "batched_cells_per_wave": g}
             for k in tables for g in GROUPS for w in GROUP_WAVES for b, t in GROUP_LU]
    return lane + group


def _time(cand: dict, integrator: str, form: str, st, m, dt: float, rtol: float, atol: float,
          calls: int, chained: bool = False):
    """(kernel ms of the last call, final state (ncells, N+1) as T and mass fractions)."""
    import torch
    from .reactor import ReactorBase
    dev = torch.device("cuda:0")
    R = ReactorBase.create(integrator)
    R.init(2, st.ncells, mechanism=cand["kernels"], rtol=rtol, atol=atol,
           **{k: v for k, v in cand.items() if k != "kernels"})
    ns = m.n_species
    ms = 0.0
    if form == "box":
        n = st.ncells
        shape = (1, 1, n)

        def soa(a, nc):
            return torch.from_numpy(np.ascontiguousarray(a.T.reshape((nc,) + shape))).to(dev)

        rY0, T0, rE0 = soa(st.rY, ns), soa(st.T[:, None], 1), soa(st.rE[:, None], 1)
        rYs, rEs, FC = torch.zeros_like(rY0), torch.zeros_like(rE0), torch.zeros_like(T0)
        rY, T, rE = rY0.clone(), T0.clone(), rE0.clone()  # the same arrays each call: cost order
        for _ in range(calls):
            if not chained:
                rY.copy_(rY0)
                T.copy_(T0)
                rE.copy_(rE0)
            R.react(rY, rYs, T, rE, rEs, FC, dt)
            ms = R.stats()["kernel_ms"]
        out = np.concatenate([rY.reshape(ns, -1).T.cpu().numpy(), T.reshape(-1, 1).cpu().numpy()], axis=1)
    else:
        y0 = torch.from_numpy(st.y_1d()).to(dev)
        x0 = torch.from_numpy(st.rE.copy()).to(dev)
        ys = torch.zeros(st.ncells * ns, dtype=torch.float64, device=dev)
        xs = torch.zeros(st.ncells, dtype=torch.float64, device=dev)
        y, x = y0.clone(), x0.clone()
        for _ in range(calls):
            if not chained:
                y.copy_(y0)
                x.copy_(x0)
            R.react_1d(y, ys, x, xs, dt)
            ms = R.stats()["kernel_ms"]
        out = y.reshape(st.ncells, ns + 1).cpu().numpy()
    R.close()
    Y = out[:, :ns] / out[:, :ns].sum(1, keepdims=True)
    return ms, np.concatenate([Y, out[:, ns:]], axis=1)


def _agrees(a, b) -> bool:
    """Per-component |a - b| <= 1e-6 |b| + 1e-10 on mass fractions and T."""
    return bool(np.all(np.abs(a - b) <= 1e-6 * np.abs(b) + 1e-10))


def tune(mech: str, integrator: str, form: str = "box", ncells: int = 262144, dt: float = 1e-5,
         rtol: float = 1e-10, atol: float = 1e-16, calls: int = 2, out: str | None = None,
         log=print, chained: bool = False, groups: bool = True) -> dict:
    from .mechgen.__main__ import load
    from .throughput import workload
    st = workload(mech, ncells)
    m = load(mech)
    from .reactor import ReactorError
    results, ref = [], None
    for cand in candidates(mech, integrator, form):
        if not groups and cand.get("batched_cells_per_wave"):
            continue
        # every launch shape integrates each cell with the same arithmetic: a
        # candidate whose cells fail or whose result leaves the criterion
        # around the first candidate's is not eligible
        try:
            ms, res = _time(cand, integrator, form, st, m, dt, rtol, atol, calls, chained)
        except ReactorError as e:
            results.append(dict(cand, kernel_ms=None, rejected=str(e)))
            log(f"[autotune] {mech} {integrator} {form} {cand}: rejected ({e})")
            continue
        if ref is None:
            ref = res
        ok = _agrees(res, ref)
        results.append(dict(cand, kernel_ms=ms) if ok else dict(cand, kernel_ms=None, rejected="result differs"))
        log(f"[autotune] {mech} {integrator} {form} {cand}: {ms:.1f} ms" + ("" if ok else " rejected: result differs"))
    eligible = [r for r in results if r["kernel_ms"] is not None]
    if not eligible:
        raise RuntimeError(f"no candidate of {mech} {integrator} {form} integrated every cell")
    best = min(eligible, key=lambda r: r["kernel_ms"])
    entry = dict({k: best[k] for k in best if k != "kernel_ms"}, kernel_ms=best["kernel_ms"], cells=st.ncells,
                 dt=dt, rtol=rtol, atol=atol, calls=calls, protocol="chained" if chained else "replay",
                 candidates=results)
    a = arch()
    path = out or table_path(a)
    table = {}
    if os.path.exists(path):
        with open(path) as f:
            table = json.load(f)
    table[key(mech, integrator, form)] = entry
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return entry


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--mech", default="drm19")
    ap.add_argument("--integrator", default="ReactorBDFAdaptiveMixed")
    ap.add_argument("--form", choices=["box", "1d"], default="box")
    ap.add_argument("--cells", type=int, default=262144)
    ap.add_argument("--dt", type=float, default=1e-5)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--atol", type=float, default=1e-16)
    ap.add_argument("--calls", type=int, default=2)
    ap.add_argument("--out", default=None)
    ap.add_argument("--protocol", choices=["replay", "chained"], default="replay")
    ap.add_argument("--no-groups", action="store_true", help="skip the group-kernel candidates")
    a = ap.parse_args(argv)
    e = tune(a.mech, a.integrator, a.form, a.cells, a.dt, a.rtol, a.atol, a.calls, a.out,
             log=lambda *x: print(*x, file=sys.stderr, flush=True), chained=a.protocol == "chained",
             groups=not a.no_groups)
    print(json.dumps({k: v for k, v in e.items() if k != "candidates"}))


if __name__ == "__main__":
    main()
