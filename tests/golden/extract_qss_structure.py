"""Write golden fixtures of the quasi-steady-state (QSS) systems of the
reference's generated QSS mechanism headers.

``Mechanisms/<mech>/mechanism.H`` of a QSS mechanism solves the QSS species'
balances in ``comp_sc_qss`` from the rates ``qf_co[i]`` / ``qr_co[i]`` that
``comp_qss_coeff`` forms per QSS reaction (``// reaction <yaml index>``
comments).  This script reads both functions as text (nothing is compiled or
imported) and writes, per mechanism, ``<mech>_qss_structure.json`` with

* ``reactions``: qf_co / qr_co index -> YAML reaction index;
* ``species``: per QSS species (its name), the terms of its balance as
  ``[part, other, direction, reaction, coefficient]`` with part "num"
  (production), "denom" (consumption) or "coupling" (other = the coupled QSS
  species), direction "f" (qf_co) or "r" (qr_co);
* ``epsilon``: the regularisation constant.

Run from the repository root, where ``/root/reference`` exists:

    python tests/golden/extract_qss_structure.py CH4_lean_qss heptane_lu_qss

The fixtures are data (the reference's QSS structure); the tests compare the
balances the project's generator derives from ``qssa.yaml`` against them.
"""
from __future__ import annotations

import json
import os
import re
import sys

REF = "/root/reference/Mechanisms"
HERE = os.path.dirname(os.path.abspath(__file__))


def _function(text: str, name: str) -> str:
    start = re.search(r"^" + name + r"\(", text, re.M).start()
    depth, i = 0, text.index("{", start)
    for j in range(i, len(text)):
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
            if depth == 0:
                return text[i:j + 1]
    raise ValueError(name)


def _terms(expr: str):
    out = []
    for sign, coef, d, idx in re.findall(r"([+-])\s*(?:([0-9.]+)\s*\*\s*)?(qf_co|qr_co)\[(\d+)\]", expr):
        c = float(coef) if coef else 1.0
        out.append(("f" if d == "qf_co" else "r", int(idx), c if sign == "+" else -c))
    return out


def extract(mech: str) -> dict:
    with open(os.path.join(REF, mech, "mechanism.H")) as f:
        text = f.read()
    coeff = _function(text, "comp_qss_coeff")
    rx = {}
    for m in re.finditer(r"//\s*reaction (\d+):[^\n]*\n(.*?)\n\s*}", coeff, re.S):
        q = re.search(r"qf\[(\d+)\]\s*=", m.group(2))
        rx[int(q.group(1))] = int(m.group(1))
    body = _function(text, "comp_sc_qss")
    eps = float(re.search(r"epsilon\s*=\s*([0-9.eE+-]+)", body).group(1))
    stmts = [s.strip() for s in body.replace("\n", " ").split(";")]
    # sanitised variable prefix of each species, from "// QSSA species i: name"
    names = {}
    for m in re.finditer(r"QSSA species (\d+):\s*(?://\s*)?([^\s/]+)\s+amrex::Real\s+(\w+)_num\b", body):
        names[m.group(3)] = m.group(2)
    species = {v: [] for v in names.values()}
    for s in stmts:
        m = re.search(r"amrex::Real\s+(\w+)\s*=\s*(.*)$", s)
        if not m:
            continue
        var, expr = m.group(1), m.group(2)
        if var.endswith("_num") and var[:-4] in names:
            for d, i, c in _terms(expr):
                species[names[var[:-4]]].append(["num", None, d, rx[i], c])
        elif var.endswith("_denom") and var[:-6] in names:
            for d, i, c in _terms(expr):
                species[names[var[:-6]]].append(["denom", None, d, rx[i], -c])
        else:
            dm = re.search(r"/\s*(\w+)_denom\s*$", expr)
            if not dm or dm.group(1) not in names or not var.startswith(dm.group(1) + "_"):
                continue
            other = var[len(dm.group(1)) + 1:]
            if other not in names:
                continue
            for d, i, c in _terms(expr):
                species[names[dm.group(1)]].append(["coupling", names[other], d, rx[i], c])
    for v in species.values():
        v.sort(key=lambda t: (t[0], t[1] or "", t[2], t[3], t[4]))
    return {"mechanism": mech, "epsilon": eps, "reactions": [rx[i] for i in sorted(rx)],
            "species": species}


if __name__ == "__main__":
    for mech in sys.argv[1:] or ["CH4_lean_qss", "heptane_lu_qss"]:
        d = extract(mech)
        with open(os.path.join(HERE, f"{mech}_qss_structure.json"), "w") as f:
            json.dump(d, f, indent=1)
        print(mech, len(d["reactions"]), "QSS reactions,",
              sum(len(v) for v in d["species"].values()), "terms")
