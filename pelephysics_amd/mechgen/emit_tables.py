"""Emit the constant-memory tables of one mechanism for the table-driven kernels.

The table-driven device code (pelephysics_amd/csrc/chem_table.h) walks these
tables with a uniform (wavefront-wide) index, so every parameter is a scalar
load and the loop body is the same few hundred instructions for all
reactions, instead of one unrolled block per reaction.

Reaction order: elementary, then three-body, then falloff (Troe, then
Lindemann), each group contiguous so the kernels run one loop per group with
no per-reaction branching.  For every reaction:

* ``rxA/rxB/rxE``: k = A T^b exp(-E/T) (SI-mol units; high-pressure for falloff);
* ``rxR[3]``/``rxP[3]``: reactant/product species, repeated for coefficient 2
  or 3, -1 when unused (products only matter for reversible reactions);
* ``rxDelta``: sum(nu') - sum(nu''), the power of refC in the reverse rate;
* ``rxNet[6]``/``rxNu[6]``: net stoichiometry (species, coefficient), -1 unused;
* ``rxEff0/rxEffN`` into ``effK``/``effV``: third-body enhancement (eff - 1)
  per listed species (default efficiency 1);
* falloff: ``foA/foB/foE`` (low-pressure rate, SI-mol) and Troe parameters
  ``troe[4]`` with ``troeT2`` flag;
* ``rxIndex``: position of the reaction in the YAML file (progress-rate and
  sensitivity outputs are reported in YAML order);
* ``spPtr/spRx/spNu``: the same net stoichiometry by species (compressed rows,
  reactions in table order), for kernels that form each species' production
  rate in its own lane.

QSS mechanisms (``NQ`` > 0) index their quasi-steady-state species NS..NS+NQ-1
in ``rxR``/``rxP`` and the thermo tables; ``rxNet`` and ``spPtr`` keep the
transported species only.  Their extra tables:

* ``rxFwdOff``: 1 where the QSS reduction removed the forward direction;
* ``qRx``: the reactions with a QSS species (table order), and per such
  reaction ``qTermPtr`` into the balance terms ``tSlot`` (entry of the QSS
  matrix, or -1 - q for the right-hand side of species q), ``tMul`` (signed
  coefficient), ``tDir`` (0 forward, 1 reverse rate), ``tRow``/``tCol`` (the
  term's species and its coupled species, -1 for production);
* the QSS matrix in ``NSLOT`` entries (``slotInit``: -eps on the structural
  entries, 0 on fill-in) and its elimination without pivoting in a
  fill-reducing order: ``luA/luB/luC`` (A[a] /= A[b] when c < 0, else
  A[a] -= A[b] A[c]), ``fsI/fsS/fsP`` (b[i] -= A[s] b[p]) and, per pivot in
  reverse order, ``bsRow/bsDiag/bsPtr`` with ``bsSlot/bsCol``
  (x[row] = (b[row] - sum A[slot] x[col]) / A[diag]).
"""
from __future__ import annotations

from typing import List

from .emit_hip import _f
from .model import Mechanism, Reaction

MAXR = 3
MAXNET = 6


def ordered(m: Mechanism) -> List[int]:
    el = [j for j, r in enumerate(m.reactions) if r.kind == "elementary"]
    tb = [j for j, r in enumerate(m.reactions) if r.kind == "three-body"]
    ft = [j for j, r in enumerate(m.reactions) if r.kind == "falloff" and r.troe is not None]
    fl = [j for j, r in enumerate(m.reactions) if r.kind == "falloff" and r.troe is None]
    return el + tb + ft + fl


def _expand(stoich) -> List[int]:
    out = []
    for k, nu in sorted(stoich.items()):
        n = int(round(nu))
        if abs(nu - n) > 1e-12 or n < 1:
            raise ValueError("non-integer stoichiometry")
        out += [k] * n
    if len(out) > MAXR:
        raise ValueError("more than 3 reactant/product molecules")
    return out + [-1] * (MAXR - len(out))


class _Tables:
    """Members of the one constant-memory struct of a mechanism: a single base
    address serves every table (one scalar register pair in the kernels)."""

    def __init__(self):
        self.items = []

    def __call__(self, name: str, ctype: str, vals):
        vals = list(vals)
        if not vals:
            vals = ["0" if ctype != "double" else "0.0"]
        self.items.append((name, ctype, vals))

    def emit(self, w, sname: str):
        w(f"struct {sname} {{")
        for name, ctype, vals in self.items:
            w(f"  {ctype} {name}[{len(vals)}];")
        w("};")
        w("")
        w(f"static __constant__ {sname} kTab = {{")
        for name, _, vals in self.items:
            w(f"    /* {name} */ {{{', '.join(vals)}}},")
        w("};")


def qss_elimination(m: Mechanism):
    """Slots of the QSS matrix, their initial values and the elimination
    program (no pivoting, greedy minimum-degree order on the symmetrised
    pattern)."""
    from .model import QSS_EPS
    nq = m.n_qss
    terms = m.qss_terms()
    struct = {(q, q) for q in range(nq)} | {(r, c) for r, c, _, _, _ in terms if c >= 0}
    pat = set(struct)
    left = set(range(nq))
    order = []
    while left:
        def degree(p):
            return len({i for i in left if i != p and ((i, p) in pat or (p, i) in pat)})
        p = min(sorted(left), key=degree)
        order.append(p)
        left.discard(p)
        rows = [i for i in sorted(left) if (i, p) in pat]
        cols = [j for j in sorted(left) if (p, j) in pat]
        for i in rows:
            for j in cols:
                pat.add((i, j))
    slots = sorted(pat, key=lambda e: (order.index(e[0]), order.index(e[1])))
    slot = {e: k for k, e in enumerate(slots)}
    init = [(-QSS_EPS if e in struct else 0.0) for e in slots]
    lu, fs, bs = [], [], []
    done = []
    for p in order:
        rest = [q for q in order if q not in done and q != p]
        rows = [i for i in rest if (i, p) in pat]
        cols = [j for j in rest if (p, j) in pat]
        for i in rows:
            lu.append((slot[(i, p)], slot[(p, p)], -1))
            for j in cols:
                lu.append((slot[(i, j)], slot[(i, p)], slot[(p, j)]))
            fs.append((i, slot[(i, p)], p))
        done.append(p)
    for p in reversed(order):
        later = order[order.index(p) + 1:]
        bs.append((p, slot[(p, p)], [(slot[(p, j)], j) for j in later if (p, j) in pat]))
    return terms, slot, init, lu, fs, bs


def emit(m: Mechanism, source: str) -> str:
    order = ordered(m)
    R = [m.reactions[j] for j in order]
    n_el = sum(r.kind == "elementary" for r in R)
    n_tb = sum(r.kind == "three-body" for r in R)
    n_troe = sum(r.kind == "falloff" and r.troe is not None for r in R)
    n_lind = sum(r.kind == "falloff" and r.troe is None for r in R)
    ns = m.n_species
    nq = m.n_qss
    L = []
    w = L.append
    w(f"// Generated by pelephysics_amd.mechgen.emit_tables from {source}. Do not edit.")
    w("#pragma once")
    w('#include "../ppc_common.h"')
    w("")
    w("namespace ppc {")
    w("namespace tab {")
    w(f"namespace {m.name}_t {{")
    w("")
    tab = _Tables()
    # thermo
    lo = [c for s in m.all_species for c in s.low]
    hi = [c for s in m.all_species for c in s.high]
    tab("nasaLo", "double", (_f(x) for x in lo))
    tab("nasaHi", "double", (_f(x) for x in hi))
    tab("tmid", "double", (_f(s.tmid) for s in m.all_species))
    tab("mw", "double", (_f(s.mw) for s in m.species))
    tab("imw", "double", (_f(s.imw) for s in m.species))
    # reactions
    tab("rxA", "double", (_f(r.rate.A) for r in R))
    tab("rxB", "double", (_f(r.rate.b) for r in R))
    tab("rxE", "double", (_f(r.rate.EaR) for r in R))
    rr, pp, net, nu, delta, rev = [], [], [], [], [], []
    for r in R:
        rr += _expand(r.reactants)
        pp += _expand(r.products) if r.reversible else [-1] * MAXR
        ns_ = sorted((k, v) for k, v in r.net_stoich().items() if k < ns)
        if len(ns_) > MAXNET:
            raise ValueError("more than 6 species in one reaction")
        net += [k for k, _ in ns_] + [-1] * (MAXNET - len(ns_))
        nu += [_f(v) for _, v in ns_] + ["0.0"] * (MAXNET - len(ns_))
        d = r.delta_order()
        delta.append(str(int(round(d))))
        rev.append("1" if r.reversible else "0")
    tab("rxR", "int", map(str, rr))
    tab("rxP", "int", map(str, pp))
    tab("rxNet", "int", map(str, net))
    tab("rxNu", "double", nu)
    tab("rxDelta", "int", delta)
    tab("rxRev", "int", rev)
    eff0, effn, effk, effv = [], [], [], []
    for r in R:
        eff0.append(str(len(effk)))
        cnt = 0
        if r.has_third_body:
            if r.default_eff != 1.0:
                raise ValueError("default efficiency other than 1 is not supported")
            for k, e in sorted(r.efficiencies.items()):
                if e != 1.0:
                    effk.append(str(k))
                    effv.append(_f(e - 1.0))
                    cnt += 1
        effn.append(str(cnt))
    tab("rxEff0", "int", eff0)
    tab("rxEffN", "int", effn)
    tab("effK", "int", effk)
    tab("effV", "double", effv)
    fo = [r for r in R if r.kind == "falloff"]
    tab("foA", "double", (_f(r.low.A) for r in fo))
    tab("foB", "double", (_f(r.low.b) for r in fo))
    tab("foE", "double", (_f(r.low.EaR) for r in fo))
    tr = [r for r in R if r.kind == "falloff" and r.troe is not None]
    tv = []
    for r in tr:
        t = list(r.troe) + [0.0] * (4 - len(r.troe))
        tv += [_f(x) for x in t]
    tab("troe", "double", tv)
    tab("troeT2", "int", ("1" if len(r.troe) == 4 else "0" for r in tr))
    tab("rxIndex", "int", (str(j) for j in order))
    rows = [[] for _ in range(ns)]
    for jj, r in enumerate(R):
        for k, v in sorted(r.net_stoich().items()):
            if k < ns:
                rows[k].append((jj, v))
    ptr, srx, snu = [0], [], []
    for k in range(ns):
        srx += [str(jj) for jj, _ in rows[k]]
        snu += [_f(v) for _, v in rows[k]]
        ptr.append(len(srx))
    tab("spPtr", "int", map(str, ptr))
    tab("spRx", "int", srx)
    tab("spNu", "double", snu)
    qss_counts = {"NQR": 0, "NSLOT": 1, "NLU": 0, "NFS": 0}
    if nq:
        terms, slot, init, lu, fs, bs = qss_elimination(m)
        pos = {j: jj for jj, j in enumerate(order)}
        tab("rxFwdOff", "int", ("1" if r.fwd_removed else "0" for r in R))
        qrx = sorted(pos[j] for j in m.qss_reactions())
        by_rx = {jj: [] for jj in qrx}
        for q, c, j, d, cf in terms:
            if c < 0:
                sl, mul = -1 - q, cf
            elif c == q:
                sl, mul = slot[(q, q)], cf
            else:
                sl, mul = slot[(q, c)], -cf
            by_rx[pos[j]].append((sl, mul, d, q, c))
        tp, tsl, tmul, tdir, trow, tcol = [0], [], [], [], [], []
        for jj in qrx:
            for sl, mul, d, q, c in by_rx[jj]:
                tsl.append(str(sl))
                tmul.append(_f(mul))
                tdir.append(str(d))
                trow.append(str(q))
                tcol.append(str(c))
            tp.append(len(tsl))
        tab("qRx", "int", map(str, qrx))
        tab("qTermPtr", "int", map(str, tp))
        tab("tSlot", "int", tsl)
        tab("tMul", "double", tmul)
        tab("tDir", "int", tdir)
        tab("tRow", "int", trow)
        tab("tCol", "int", tcol)
        tab("slotInit", "double", (_f(v) for v in init))
        tab("luA", "int", (str(a) for a, _, _ in lu))
        tab("luB", "int", (str(b) for _, b, _ in lu))
        tab("luC", "int", (str(c) for _, _, c in lu))
        tab("fsI", "int", (str(i) for i, _, _ in fs))
        tab("fsS", "int", (str(sl) for _, sl, _ in fs))
        tab("fsP", "int", (str(p) for _, _, p in fs))
        bptr, bsl, bcol = [0], [], []
        for _, _, ents in bs:
            bsl += [str(sl) for sl, _ in ents]
            bcol += [str(c) for _, c in ents]
            bptr.append(len(bsl))
        tab("bsRow", "int", (str(p) for p, _, _ in bs))
        tab("bsDiag", "int", (str(d) for _, d, _ in bs))
        tab("bsPtr", "int", map(str, bptr))
        tab("bsSlot", "int", bsl)
        tab("bsCol", "int", bcol)
        qss_counts = {"NQR": len(qrx), "NSLOT": len(slot), "NLU": len(lu), "NFS": len(fs)}
    tab.emit(w, "Tables")
    w("")
    w("}  // namespace")
    w("")
    w(f"struct {m.name} {{")
    w(f"  static constexpr int NS = {ns};")
    w(f"  static constexpr int NR = {m.n_reactions};")
    w(f"  static constexpr int NEQ = {ns + 1};")
    w(f"  static constexpr int NQ = {nq}, NX = {ns + nq};")
    w("  static constexpr int " + ", ".join(f"{k} = {v}" for k, v in qss_counts.items()) + ";")
    w(f"  static constexpr int N_EL = {n_el}, N_TB = {n_tb}, N_TROE = {n_troe}, N_LIND = {n_lind};")
    w(f"  static constexpr int MAXR = {MAXR}, MAXNET = {MAXNET};")
    for nm in ("nasaLo", "nasaHi", "tmid", "mw", "imw", "rxA", "rxB", "rxE", "rxNu", "effV", "foA",
               "foB", "foE", "troe", "spNu"):
        w(f"  PPC_DI static double {nm}_(int i) {{ return {m.name}_t::kTab.{nm}[i]; }}")
    for nm in ("rxR", "rxP", "rxNet", "rxDelta", "rxRev", "rxEff0", "rxEffN", "effK", "troeT2",
               "rxIndex", "spPtr", "spRx"):
        w(f"  PPC_DI static int {nm}_(int i) {{ return {m.name}_t::kTab.{nm}[i]; }}")
    if nq:
        for nm in ("tMul", "slotInit"):
            w(f"  PPC_DI static double {nm}_(int i) {{ return {m.name}_t::kTab.{nm}[i]; }}")
        for nm in ("rxFwdOff", "qRx", "qTermPtr", "tSlot", "tDir", "tRow", "tCol", "luA", "luB", "luC", "fsI",
                   "fsS", "fsP", "bsRow", "bsDiag", "bsPtr", "bsSlot", "bsCol"):
            w(f"  PPC_DI static int {nm}_(int i) {{ return {m.name}_t::kTab.{nm}[i]; }}")
    w("};")
    w("")
    w("}  // namespace tab")
    w("}  // namespace ppc")
    return "\n".join(L) + "\n"
