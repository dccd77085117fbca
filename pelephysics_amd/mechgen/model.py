"""Chemical mechanism model: Cantera-format YAML -> normalised parameter tables.

The tables reproduce the constants that the reference's generated mechanism
headers bake into code (e.g. ``Mechanisms/drm19/mechanism.H:4421-5667`` for
``productionRate`` and ``:10307-15357`` for ``aJacobian``):

* species order = the YAML phase species list (``mechanism.H:16-36`` ``*_ID``);
* molecular weights from element weights, rounded to 6 decimals, and inverse
  weights rounded to 16 decimals (``mechanism.H:101-275``);
* rate constants in SI-mol units (m^3/mol), i.e. the CGS pre-exponential times
  1e-6 per concentration order (``Support/ceptr/ceptr/production.py:187-233``);
* activation temperatures Ea/R with Ea in cal/mol and Rc = 1.98721558317399615845
  cal/mol/K (``Source/PhysicsConstants.H:12``; e.g. 6290 cal/mol ->
  3165.23282791169 K in ``Mechanisms/LiDryer/mechanism.H`` reaction 1);
* third-body efficiencies as (default, {species: efficiency});
* Troe parameters (A, T3, T1[, T2]) and Lindemann falloff.

The model is plain data and serialises to JSON; it is consumed by the HIP code
emitter (``emit_hip``) and by the CPU oracle (``oracle/``).
"""
from __future__ import annotations

import json
import math
import re
from dataclasses import dataclass, field, asdict
from typing import Dict, List, Optional, Tuple

# Element atomic weights (g/mol) as used by Cantera 2.6/3.x for the elements
# present in the mechanisms of interest.  They reproduce the molecular weights
# printed in the reference headers (e.g. CH4 16.043, AR 39.95).
ELEMENT_WEIGHTS = {
    "H": 1.008,
    "C": 12.011,
    "N": 14.007,
    "O": 15.999,
    "Ar": 39.95,
    "AR": 39.95,
    "He": 4.002602,
    "HE": 4.002602,
}

# cal/mol -> K: Ea / Rc with Rc in cal/(mol K) (Source/PhysicsConstants.H:12).
RUC = 1.98721558317399615845
CAL_PER_MOL_TO_K = 1.0 / RUC

# Physical constants of the reference (Source/PhysicsConstants.H:10-13).
RU = 8.31446261815324e7  # erg/(mol K)
PATM = 1.01325e6  # dyn/cm^2


@dataclass
class Arrhenius:
    """k = A T^b exp(-EaR / T), A in SI-mol units."""

    A: float
    b: float
    EaR: float

    @property
    def lnA(self) -> float:
        return math.log(self.A)


@dataclass
class Reaction:
    index: int  # index in the YAML file (reference "reaction <idx>" comments)
    equation: str
    reactants: Dict[int, float]  # species index -> stoichiometric coefficient
    products: Dict[int, float]
    reversible: bool
    kind: str  # "elementary" | "three-body" | "falloff"
    rate: Arrhenius  # high-pressure rate for falloff
    low: Optional[Arrhenius] = None  # low-pressure rate (falloff only)
    troe: Optional[List[float]] = None  # [A, T3, T1] or [A, T3, T1, T2]
    default_eff: float = 1.0
    efficiencies: Dict[int, float] = field(default_factory=dict)
    # forward direction dropped by the QSS reduction (qssa.yaml
    # forward_to_remove_idx): the reaction runs in reverse only
    fwd_removed: bool = False

    @property
    def has_third_body(self) -> bool:
        return self.kind in ("three-body", "falloff")

    def net_stoich(self) -> Dict[int, float]:
        nu: Dict[int, float] = {}
        for k, v in self.reactants.items():
            nu[k] = nu.get(k, 0.0) - v
        for k, v in self.products.items():
            nu[k] = nu.get(k, 0.0) + v
        return {k: v for k, v in nu.items() if v != 0.0}

    def delta_order(self) -> float:
        """sum(reactant coefficients) - sum(product coefficients)."""
        return sum(self.reactants.values()) - sum(self.products.values())


@dataclass
class Species:
    name: str
    composition: Dict[str, float]
    mw: float
    imw: float
    tmid: float
    low: List[float]  # NASA7 coefficients a1..a7 for T < tmid
    high: List[float]  # NASA7 coefficients a1..a7 for T >= tmid


# Regularisation of the quasi-steady-state balances: every structural entry of
# the QSS system (diagonal, coupling, right-hand side) carries it, as the
# reference's generated comp_sc_qss does (epsilon = 1e-12,
# Support/ceptr/ceptr/qssa_converter.py).
QSS_EPS = 1e-12


@dataclass
class Mechanism:
    """species: the transported species (the state); qss_species: species in
    quasi-steady state, indexed n_species + q in reaction stoichiometry."""

    name: str
    elements: List[str]
    species: List[Species]
    reactions: List[Reaction]
    qss_species: List[Species] = field(default_factory=list)

    @property
    def n_species(self) -> int:
        return len(self.species)

    @property
    def n_qss(self) -> int:
        return len(self.qss_species)

    @property
    def all_species(self) -> List[Species]:
        return self.species + self.qss_species

    def qss_terms(self) -> List[Tuple[int, int, int, int, float]]:
        """The quasi-steady-state balance of each QSS species q as terms
        (q, col, reaction, direction, coefficient): direction 0 takes the
        forward rate and 1 the reverse rate of the reaction with the QSS
        concentration factored out (qf_co / qr_co); col = -1 adds
        coefficient x rate to the production P_q, col = q adds it to the
        consumption D_q, another col c to the coupling C_qc.  The balance is
        (D_q - eps) x_q - sum_c (C_qc + eps) x_c = P_q + eps.

        The rules restate Support/ceptr/ceptr/qssa_converter.py:960-1250: a
        reaction with one QSS species contributes its consumption (rate x 1)
        and production (stoichiometric coefficient x rate); with two, the
        rate in the direction that turns one into the other couples them; a
        reaction whose forward direction was removed and whose QSS species
        are all reactants contributes only its reverse production, and an
        irreversible reaction producing two QSS species contributes nothing."""
        NS = self.n_species
        terms = []
        for q in range(self.n_qss):
            s = NS + q
            for j, r in enumerate(self.reactions):
                inR, inP = s in r.reactants, s in r.products
                if not (inR or inP):
                    continue
                if inR and inP:
                    raise ValueError(f"QSS species on both sides of {r.equation!r}")
                Q = sorted({k for k in list(r.reactants) + list(r.products) if k >= NS})
                nuR, nuP = r.reactants.get(s, 0.0), r.products.get(s, 0.0)
                if len(Q) == 1:
                    if inR:
                        terms.append((q, q, j, 0, 1.0))
                        if r.reversible:
                            terms.append((q, -1, j, 1, nuR))
                    else:
                        terms.append((q, -1, j, 0, nuP))
                        if r.reversible:
                            terms.append((q, q, j, 1, 1.0))
                    continue
                if len(Q) > 2:
                    raise ValueError(f"more than two QSS species in {r.equation!r}")
                o = Q[0] if Q[1] == s else Q[1]
                if not any(k in r.products for k in Q) and r.fwd_removed:
                    if r.reversible:
                        terms.append((q, -1, j, 1, nuR))
                elif inR:
                    terms.append((q, q, j, 0, 1.0))
                    if r.reversible:
                        terms.append((q, o - NS, j, 1, nuR))
                elif o in r.reactants:
                    terms.append((q, o - NS, j, 0, nuP))
                    if r.reversible:
                        terms.append((q, q, j, 1, 1.0))
        return terms

    def qss_reactions(self) -> List[int]:
        """Reactions (YAML order) with a QSS species on either side."""
        NS = self.n_species
        return [j for j, r in enumerate(self.reactions)
                if any(k >= NS for k in list(r.reactants) + list(r.products))]

    @property
    def n_reactions(self) -> int:
        return len(self.reactions)

    def species_index(self, name: str) -> int:
        """Index of a species by name; case-insensitive when there is no exact
        match (heptane_lu_88sk spells its species in lower case)."""
        for i, s in enumerate(self.species):
            if s.name == name:
                return i
        for i, s in enumerate(self.species):
            if s.name.lower() == name.lower():
                return i
        raise KeyError(name)

    # ------------------------------------------------------------------ JSON
    def to_json(self) -> str:
        d = {
            "name": self.name,
            "elements": self.elements,
            "species": [asdict(s) for s in self.species],
            "reactions": [],
        }
        if self.qss_species:
            d["qss_species"] = [asdict(s) for s in self.qss_species]
        for r in self.reactions:
            rd = asdict(r)
            if not rd["fwd_removed"]:
                del rd["fwd_removed"]
            # JSON object keys must be strings
            for key in ("reactants", "products", "efficiencies"):
                rd[key] = {str(k): v for k, v in rd[key].items()}
            d["reactions"].append(rd)
        return json.dumps(d, indent=1, sort_keys=False)

    @staticmethod
    def from_json(text: str) -> "Mechanism":
        d = json.loads(text)
        species = [Species(**s) for s in d["species"]]
        reactions = []
        for rd in d["reactions"]:
            rd = dict(rd)
            for key in ("reactants", "products", "efficiencies"):
                rd[key] = {int(k): float(v) for k, v in rd[key].items()}
            rd["rate"] = Arrhenius(**rd["rate"])
            rd["low"] = Arrhenius(**rd["low"]) if rd["low"] is not None else None
            reactions.append(Reaction(**rd))
        qss = [Species(**s) for s in d.get("qss_species", [])]
        return Mechanism(d["name"], d["elements"], species, reactions, qss)


# ---------------------------------------------------------------- YAML input
_COLLIDER_RE = re.compile(r"\(\+\s*([^)\s]+)\s*\)")


def _parse_side(text: str, names: Dict[str, int]) -> Tuple[Dict[int, float], bool, Optional[str]]:
    """Parse one side of an equation. Returns (stoich, has_plus_M, falloff_collider)."""
    collider = None
    m = _COLLIDER_RE.search(text)
    if m:
        collider = m.group(1)
        text = _COLLIDER_RE.sub(" ", text)
    stoich: Dict[int, float] = {}
    third = False
    for tok in text.split(" + "):
        tok = tok.strip()
        if not tok:
            continue
        parts = tok.split()
        if len(parts) == 2:
            coef, sp = float(parts[0]), parts[1]
        elif len(parts) == 1:
            coef, sp = 1.0, parts[0]
            mm = re.match(r"^(\d+(?:\.\d+)?)([A-Za-z].*)$", sp)
            if mm and sp not in names:
                coef, sp = float(mm.group(1)), mm.group(2)
        else:
            raise ValueError(f"cannot parse stoichiometric term {tok!r}")
        if sp == "M":
            third = True
            continue
        if sp not in names:
            raise KeyError(f"unknown species {sp!r} in {text!r}")
        k = names[sp]
        stoich[k] = stoich.get(k, 0.0) + coef
    return stoich, third, collider


# Units of a Cantera YAML file: (concentration unit in mol/m^3, activation
# energy unit x K per Ea/R).  A file without a "units" block is in Cantera's
# defaults (m, kmol, J/kmol), as the QSS converter writes qssa.yaml; energies
# in joules convert to cal/mol with 4.184 J/cal before the division by Rc, as
# the reference's generator does (Mechanisms/CH4_lean_qss/mechanism.H reaction
# 0: Ea = 5.14632e6 J/kmol -> 618.956498939805 K).
def _unit_system(units: dict):
    length = units.get("length", "m")
    quantity = units.get("quantity", "kmol")
    energy = units.get("activation-energy", "J/kmol")
    conc = {("cm", "mol"): 1e6, ("m", "kmol"): 1e3, ("m", "mol"): 1.0}.get((length, quantity))
    rgas = {"cal/mol": RUC, "J/kmol": 4184.0 * RUC, "J/mol": 4.184 * RUC}.get(energy)
    if conc is None or rgas is None:
        raise ValueError(f"unsupported units {units}")
    return conc, rgas


def _arrhenius(d: dict, order: float, units=(1e6, RUC)) -> Arrhenius:
    """Pre-exponential of a rate of the given concentration order in the
    file's units (default CGS-mol, cal/mol) -> SI-mol, activation energy -> K."""
    conc, rgas = units
    A = float(d["A"]) * conc ** (-(order - 1.0))
    return Arrhenius(A=A, b=float(d.get("b", 0.0)), EaR=float(d.get("Ea", 0.0)) / rgas)


def load_yaml(path: str, name: Optional[str] = None) -> Mechanism:
    import yaml

    with open(path) as f:
        doc = yaml.safe_load(f)
    units = _unit_system(doc.get("units", {}))

    phase = doc["phases"][0]
    # QSS mechanisms (qssa.yaml): transported species first, then the QSS
    # species, each in phase order
    qss_names = set(phase.get("qssa_species", []))
    sp_names: List[str] = ([n for n in phase["species"] if n not in qss_names] +
                           [n for n in phase["species"] if n in qss_names])
    n_transported = len(sp_names) - len(qss_names)
    removed = set(phase.get("forward_to_remove_idx", []))
    sp_by_name = {s["name"]: s for s in doc["species"]}
    names = {n: i for i, n in enumerate(sp_names)}

    species: List[Species] = []
    for n in sp_names:
        s = sp_by_name[n]
        comp = {k: float(v) for k, v in s["composition"].items()}
        mw = sum(ELEMENT_WEIGHTS[e] * c for e, c in comp.items())
        mw = float(f"{mw:.6f}")
        imw = float(f"{1.0 / mw:.16f}")
        th = s["thermo"]
        if th["model"] != "NASA7":
            raise ValueError(f"species {n}: only NASA7 thermo is supported")
        tr = th["temperature-ranges"]
        data = th["data"]
        if len(data) == 1:
            low = high = [float(x) for x in data[0]]
            tmid = float(tr[-1])
        else:
            low = [float(x) for x in data[0]]
            high = [float(x) for x in data[1]]
            tmid = float(tr[1])
        species.append(Species(n, comp, mw, imw, tmid, low, high))

    reactions: List[Reaction] = []
    for idx, rd in enumerate(doc["reactions"]):
        eq = rd["equation"]
        if "<=>" in eq:
            lhs, rhs = eq.split("<=>")
            rev = True
        elif "=>" in eq:
            lhs, rhs = eq.split("=>")
            rev = False
        elif "=" in eq:
            lhs, rhs = eq.split("=")
            rev = True
        else:
            raise ValueError(eq)
        if rd.get("orders") or rd.get("reverse") or rd.get("negative-A"):
            raise ValueError(f"unsupported reaction option in {eq!r}")
        reac, tb_l, col_l = _parse_side(lhs, names)
        prod, tb_r, col_r = _parse_side(rhs, names)
        rtype = rd.get("type", "elementary")
        order = sum(reac.values())
        eff = {names[k]: float(v) for k, v in rd.get("efficiencies", {}).items() if k in names}
        if any(k >= n_transported for k in eff):
            raise ValueError(f"third-body efficiency of a QSS species in {eq!r}")
        default_eff = float(rd.get("default-efficiency", 1.0))
        if rtype == "elementary":
            if tb_l or tb_r or col_l:
                raise ValueError(f"elementary reaction with third body: {eq!r}")
            r = Reaction(idx, eq, reac, prod, rev, "elementary", _arrhenius(rd["rate-constant"], order, units))
        elif rtype == "three-body":
            if not (tb_l and tb_r):
                raise ValueError(f"explicit-collider three-body reactions unsupported: {eq!r}")
            r = Reaction(idx, eq, reac, prod, rev, "three-body",
                         _arrhenius(rd["rate-constant"], order + 1.0, units),
                         default_eff=default_eff, efficiencies=eff)
        elif rtype == "falloff":
            if col_l != "M":
                raise ValueError(f"explicit-collider falloff reactions unsupported: {eq!r}")
            troe = None
            if "Troe" in rd:
                t = rd["Troe"]
                troe = [float(t["A"]), float(t["T3"]), float(t["T1"])]
                if "T2" in t:
                    troe.append(float(t["T2"]))
            elif "SRI" in rd:
                raise ValueError(f"SRI falloff unsupported: {eq!r}")
            r = Reaction(idx, eq, reac, prod, rev, "falloff",
                         _arrhenius(rd["high-P-rate-constant"], order, units),
                         low=_arrhenius(rd["low-P-rate-constant"], order + 1.0, units),
                         troe=troe, default_eff=default_eff, efficiencies=eff)
        else:
            raise ValueError(f"unsupported reaction type {rtype!r}: {eq!r}")
        r.fwd_removed = idx in removed
        if r.fwd_removed and not rev:
            raise ValueError(f"forward direction removed from irreversible {eq!r}")
        # the QSS balance is linear in the QSS concentrations
        if sum(v for k, v in reac.items() if k >= n_transported) > 1 and not r.fwd_removed:
            raise ValueError(f"QSS species quadratic in the forward rate of {eq!r}")
        if sum(v for k, v in prod.items() if k >= n_transported) > 1 and rev:
            raise ValueError(f"QSS species quadratic in the reverse rate of {eq!r}")
        reactions.append(r)

    elements = [str(e) for e in phase["elements"]]
    m = Mechanism(name or phase.get("name", "gas"), elements, species[:n_transported], reactions,
                  species[n_transported:])
    m.qss_terms()  # validates the QSS structure
    return m
