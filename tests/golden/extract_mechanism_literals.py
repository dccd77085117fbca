"""Write golden fixtures of the constants baked into the reference mechanism headers.

The reference's generated ``Mechanisms/<mech>/mechanism.H`` carries every rate
and thermodynamic constant as a C literal.  This script reads those headers as
text (nothing is compiled or imported), evaluates the literal expressions at a
fixed temperature grid and writes, per mechanism, a JSON fixture with:

* ``mw``: molecular weights (``h_global_mw``);
* ``thermo``: cp/R, h/RT and g/RT per species on the grid (functions ``cp_R``,
  ``speciesEnthalpy``, ``gibbs``);
* ``reactions``: for each ``productionRate`` block, the forward rate constant
  k_f(T), the low-pressure rate k_0(T) and the Troe F_cent(T) on the grid, the
  third-body enhancement (efficiency - 1 per species), reactant and net
  stoichiometry, reversibility and the power of ``refC`` in the reverse rate;
  for falloff blocks also ``falloff``: per grid temperature, pairs
  (log10 Pr, falloff factor) where the factor is the block's own statement
  chain from ``redP`` to the final ``Corr`` (F * F_troe for Troe blocks,
  redP / (1 + redP) for Lindemann blocks) evaluated with redP = Pr, at
  log10 Pr = -troe_c + {-2, -1, 0, 1, 2} (troe_c of the block at that T;
  Lindemann blocks at log10 Pr = -2..2).

Run from the repository root, where ``/root/reference`` exists:

    python tests/golden/extract_mechanism_literals.py LiDryer drm19

The fixtures are data (the expected outputs of the reference's constants); the
tests compare the project's YAML-derived tables against them.
"""
from __future__ import annotations

import json
import math
import os
import re
import sys

REF = "/root/reference/Mechanisms"
TGRID = [300.0, 650.0, 999.5, 1000.0, 1200.0, 1800.0, 2500.0, 3400.0]
HERE = os.path.dirname(os.path.abspath(__file__))


def _c_to_py(expr: str) -> str:
    expr = expr.replace("amrex::Real", "")
    expr = re.sub(r"(\d)\.(?!\d)", r"\1.0", expr)  # "1." -> "1.0"
    return expr


def _eval(expr: str, env: dict) -> float:
    ns = {"exp": math.exp, "log": math.log, "log10": math.log10, "pow": math.pow,
          "M_LN10": math.log(10.0), "fabs": abs}
    ns.update(env)
    return float(eval(_c_to_py(expr), {"__builtins__": {}}, ns))


def _function_body(text: str, name: str) -> str:
    m = re.search(r"\n" + re.escape(name) + r"\s*\(", text)
    if not m:
        raise KeyError(name)
    start = text.index("{", m.end())
    depth, i = 0, start
    while True:
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
            if depth == 0:
                return text[start + 1:i]
        i += 1


def _thermo_values(text: str, fname: str, nsp: int):
    body = _function_body(text, fname)
    # collect (branch, species, expr); branch = None | ("lt", Tmid) | ("ge", Tmid)
    entries = []
    branch = None
    lines = body.split("\n")
    stmt = ""
    for ln in lines:
        s = ln.strip()
        m = re.match(r"if \(T < ([\d.e+]+)\) \{", s)
        if m:
            branch = ("lt", float(m.group(1)))
            continue
        if s.startswith("} else {"):
            branch = ("ge", branch[1])
            continue
        if s == "}":
            branch = None
            continue
        if s.startswith("species[") or stmt:
            stmt += " " + s
            if stmt.rstrip().endswith(";"):
                mm = re.match(r"\s*species\[(\d+)\]\s*=\s*(.*);\s*$", stmt)
                entries.append((branch, int(mm.group(1)), mm.group(2)))
                stmt = ""
    out = []
    for T in TGRID:
        env = {"T": T, "T2": T * T, "T3": T * T * T, "T4": T * T * T * T,
               "invT": 1.0 / T, "logT": math.log(T)}
        vals = [None] * nsp
        for br, k, e in entries:
            if br is None or (br[0] == "lt" and T < br[1]) or (br[0] == "ge" and T >= br[1]):
                vals[k] = _eval(e, env)
        out.append(vals)
    return out


def _production_blocks(text: str):
    body = _function_body(text, "productionRate")
    pieces = re.split(r"\n\s*\{\s*\n\s*// reaction ", body)[1:]
    blocks = []
    for p in pieces:
        head, rest = p.split("\n", 1)
        m = re.match(r"(\d+):\s*(.*)$", head.strip())
        flat = " ".join(rest.split())
        blocks.append((int(m.group(1)), m.group(2).strip(), flat))
    return blocks


def _falloff_chain(flat: str):
    """(log10 Pr, factor) pairs on the grid from the block's statements after redP."""
    seg = flat[flat.index("redP ="):]
    stmts = re.findall(r"(?:const amrex::Real )?(\w+) =\s*(.*?);", seg)
    assert stmts[0][0] == "redP"
    chain = []
    for name, expr in stmts[1:]:
        chain.append((name, expr))
        if name == "Corr":
            break
    else:
        raise ValueError("falloff block without a final Corr statement")
    troe_c = next((e for n, e in chain if n == "troe_c"), None)
    logfc = next((e for n, e in chain if n == "logFcent"), None)
    out = []
    for T in TGRID:
        env = {"T": T, "invT": 1.0 / T, "logT": math.log(T)}
        if troe_c is not None:
            env["logFcent"] = _eval(logfc, env)
            mid = -_eval(troe_c, env)
        else:
            mid = 0.0
        pts = []
        for off in (-2.0, -1.0, 0.0, 1.0, 2.0):
            e = dict(env)
            e["redP"] = 10.0 ** (mid + off)
            for name, expr in chain:
                e[name] = _eval(expr, e)
            pts.append([mid + off, e["Corr"]])
        out.append(pts)
    return out


def extract(mech: str) -> dict:
    with open(os.path.join(REF, mech, "mechanism.H")) as f:
        text = f.read()
    m = re.search(r"const amrex::Real h_global_mw\[(\d+)\] = \{(.*?)\};", text, re.S)
    nsp = int(m.group(1))
    mw = [float(x.split("//")[0].strip().rstrip(",")) for x in m.group(2).strip().split("\n") if x.strip()]
    thermo = {fn: _thermo_values(text, fn, nsp) for fn in ("cp_R", "speciesEnthalpy", "gibbs")}
    # QSS mechanisms: reactions also read the QSS species' g/RT and
    # concentrations (g_RT_qss, sc_qss); "reactants" and "net" list the
    # transported species only, as productionRate reports them
    mq = re.search(r"amrex::Real g_RT_qss\[(\d+)\];", text)
    nqss = int(mq.group(1)) if mq else 0

    reactions = []
    for idx, eq, flat in _production_blocks(text):
        r = {"index": idx, "equation": eq}
        kf = re.search(r"k_f = (.*?);", flat).group(1)
        r["kf"] = [_eval(kf, {"T": T, "invT": 1 / T, "logT": math.log(T)}) for T in TGRID]
        mc = re.search(r"Real Corr = (.*?);", flat)
        if mc:
            ce = mc.group(1)
            base = _eval(ce, {"mixture": 0.0, "sc": [0.0] * nsp})
            eff = {}
            for k in range(nsp):
                sc = [0.0] * nsp
                sc[k] = 1.0
                v = _eval(ce, {"mixture": 0.0, "sc": sc}) - base
                if v != 0.0:
                    eff[str(k)] = v
            r["enhancement"] = eff
            r["mixture_coef"] = _eval(ce, {"mixture": 1.0, "sc": [0.0] * nsp})
        else:
            r["enhancement"] = None
        mr = re.search(r"redP = Corr / k_f \* (.*?);", flat)
        if mr:
            r["k0"] = [_eval(mr.group(1), {"T": T, "invT": 1 / T, "logT": math.log(T)}) for T in TGRID]
        else:
            r["k0"] = None
        if mr:
            r["falloff"] = _falloff_chain(flat)
        else:
            r["falloff"] = None
        mf = re.search(r"logFcent = log10\((.*?)\);", flat)
        if mf:
            r["Fcent"] = [_eval(mf.group(1), {"T": T, "invT": 1 / T, "logT": math.log(T)}) for T in TGRID]
        else:
            r["Fcent"] = None
        qf = re.search(r"qf = (.*?);", flat).group(1)
        reac = {}
        for k in re.findall(r"sc\[(\d+)\]", qf):
            reac[k] = reac.get(k, 0.0) + 1.0
        r["reactants"] = reac
        mq = re.search(r"qr = (.*?);", flat)
        qr = mq.group(1) if mq else "0.0"
        env = {"Corr": 1.0, "k_f": 1.0, "g_RT": [0.0] * nsp, "sc": [1.0] * nsp,
               "g_RT_qss": [0.0] * nqss, "sc_qss": [1.0] * nqss, "refC": 2.0, "refCinv": 0.5}
        # some irreversible reactions are emitted with a reverse rate times a
        # literal 0.0 (e.g. heptane_lu_88sk reaction 21)
        r["reversible"] = "g_RT" in qr and _eval(qr, env) != 0.0
        if r["reversible"]:
            r["refc_power"] = math.log2(_eval(qr, env))
        else:
            r["refc_power"] = None
        net = {}
        for k, sign, coef in re.findall(r"wdot\[(\d+)\] ([+-])= (?:([\d.]+) \* )?qdot;", flat):
            c = float(coef) if coef else 1.0
            net[k] = net.get(k, 0.0) + (c if sign == "+" else -c)
        r["net"] = net
        reactions.append(r)
    reactions.sort(key=lambda d: d["index"])
    return {"mechanism": mech, "source": f"Mechanisms/{mech}/mechanism.H",
            "tgrid": TGRID, "mw": mw, "thermo": thermo, "reactions": reactions}


def main(argv):
    for mech in argv or ["LiDryer", "drm19"]:
        d = extract(mech)
        out = os.path.join(HERE, f"{mech}_literals.json")
        with open(out, "w") as f:
            json.dump(d, f, indent=0)
        print(out, len(d["reactions"]), "reactions")


if __name__ == "__main__":
    main(sys.argv[1:])
