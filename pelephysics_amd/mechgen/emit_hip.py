"""Emit a HIP (gfx950) device header for one mechanism.

The emitted struct ``ppc::mech::<name>`` provides, for one cell held by one
lane, the functions the reference's generated ``mechanism.H`` provides to the
reactors (drm19 line numbers):

* ``cp_R``/``cv_R``/``h_RT``/``e_RT``/``g_RT``/``dcpR_dT``: NASA7 species
  thermodynamics (``mechanism.H:285,423,561,1002,1181,1539``);
* ``h_cp_mass``/``e_cv_mass``: mixture h and cp (or e and cv) in mass units,
  fused (``CKHBMS``/``CKCPBS`` ``:2170,1692``, ``CKUBMS``/``CKCVBS``
  ``:2452,1931``);
* ``production_rate``: molar production rates in SI-mol units
  (``productionRate`` ``:4421-5667``);
* ``progress_rates``: net rate of progress of every reaction (SI-mol), used
  by the adjoint (d f / d ln A_j = nu_j q_j);
* ``jacobian<CONSP>``: analytic d wdot / d c and d wdot / d T plus the
  temperature row, with the reference's constant-pressure convention that the
  mixture-concentration derivative of third-body terms is dropped
  (``aJacobian`` ``:10307-15357``).

NASA7 ranges are selected branch-free (per-coefficient select) so lanes of one
wavefront at different temperatures execute the same instructions.
"""
from __future__ import annotations

from typing import Dict, List

from .model import Mechanism, Reaction


def _f(x: float) -> str:
    """Shortest round-trip literal of a double."""
    r = repr(float(x))
    if r in ("inf", "-inf", "nan"):
        raise ValueError(x)
    if "e" not in r and "." not in r:
        r += ".0"
    return r


def _sel(lo: float, hi: float, flag: str = "lo") -> str:
    if lo == hi:
        return _f(lo)
    return f"({flag} ? {_f(lo)} : {_f(hi)})"


class _W:
    def __init__(self):
        self.lines: List[str] = []
        self.ind = 0

    def __call__(self, s: str = ""):
        self.lines.append(("  " * self.ind + s) if s else "")

    def block(self, head: str):
        self(head + " {")
        self.ind += 1

    def end(self, tail: str = "}"):
        self.ind -= 1
        self(tail)

    def text(self) -> str:
        return "\n".join(self.lines) + "\n"


def _thermo_groups(m: Mechanism) -> Dict[float, List[int]]:
    groups: Dict[float, List[int]] = {}
    for k, s in enumerate(m.species):
        groups.setdefault(s.tmid, []).append(k)
    return groups


def _emit_species_fn(w: _W, m: Mechanism, name: str, body_fn, doc: str, needs_log=False):
    w(f"// {doc}")
    w(f"PPC_DEV static inline void {name}(double* __restrict__ out, const double T) {{")
    w.ind += 1
    w("const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;")
    w("const double invT = 1.0 / T;")
    if needs_log:
        w("const double logT = log(T);")
    w("(void)T2; (void)T3; (void)T4; (void)invT;")
    for tmid, ks in _thermo_groups(m).items():
        w(f"{{ const bool lo = T < {_f(tmid)};")
        w.ind += 1
        for k in ks:
            s = m.species[k]
            w(f"out[{k}] = {body_fn(s)};  // {s.name}")
        w.ind -= 1
        w("}")
    w.ind -= 1
    w("}")
    w()


def _cp_expr(s) -> str:
    a = [_sel(lo, hi) for lo, hi in zip(s.low, s.high)]
    return f"{a[0]} + T * ({a[1]} + T * ({a[2]} + T * ({a[3]} + T * {a[4]})))"


def _cv_expr(s) -> str:
    return f"({_cp_expr(s)}) - 1.0"


def _h_expr(s) -> str:
    lo = [s.low[0], s.low[1] / 2, s.low[2] / 3, s.low[3] / 4, s.low[4] / 5, s.low[5]]
    hi = [s.high[0], s.high[1] / 2, s.high[2] / 3, s.high[3] / 4, s.high[4] / 5, s.high[5]]
    a = [_sel(x, y) for x, y in zip(lo, hi)]
    return f"{a[0]} + T * ({a[1]} + T * ({a[2]} + T * ({a[3]} + T * {a[4]}))) + {a[5]} * invT"


def _e_expr(s) -> str:
    return f"({_h_expr(s)}) - 1.0"


def _g_expr(s) -> str:
    # g/RT = a6/T + a1 - a7 - a1 ln T - a2/2 T - a3/6 T^2 - a4/12 T^3 - a5/20 T^4
    def coefs(c):
        return [c[5], c[0] - c[6], -c[0], -c[1] / 2, -c[2] / 6, -c[3] / 12, -c[4] / 20]
    lo, hi = coefs(s.low), coefs(s.high)
    a = [_sel(x, y) for x, y in zip(lo, hi)]
    return (f"{a[0]} * invT + {a[1]} + {a[2]} * logT + "
            f"T * ({a[3]} + T * ({a[4]} + T * ({a[5]} + T * {a[6]})))")


def _dcp_expr(s) -> str:
    lo = [s.low[1], 2 * s.low[2], 3 * s.low[3], 4 * s.low[4]]
    hi = [s.high[1], 2 * s.high[2], 3 * s.high[3], 4 * s.high[4]]
    a = [_sel(x, y) for x, y in zip(lo, hi)]
    return f"{a[0]} + T * ({a[1]} + T * ({a[2]} + T * {a[3]}))"


def _prod_term(stoich: Dict[int, float], arr: str = "sc") -> str:
    terms = []
    for k, nu in sorted(stoich.items()):
        n = int(round(nu))
        if abs(nu - n) > 1e-12 or n < 1:
            raise ValueError("non-integer stoichiometry is not supported")
        terms += [f"{arr}[{k}]"] * n
    return " * ".join(terms) if terms else "1.0"


def _dphi(stoich: Dict[int, float], k: int, arr: str = "sc") -> str:
    """d(prod c^nu)/dc_k as an expression."""
    n = int(round(stoich[k]))
    rest = dict(stoich)
    if n == 1:
        del rest[k]
        return _prod_term(rest, arr) if rest else "1.0"
    rest[k] = n - 1
    return f"{_f(float(n))} * {_prod_term(rest, arr)}"


def _kc_arg(r: Reaction) -> str:
    """sum(nu' g) - sum(nu'' g): Kc = exp(arg) * refC^(-delta) as in the reference."""
    parts = []
    net: Dict[int, float] = {}
    for k, v in r.reactants.items():
        net[k] = net.get(k, 0.0) + v
    for k, v in r.products.items():
        net[k] = net.get(k, 0.0) - v
    for k, v in sorted(net.items()):
        if v == 0:
            continue
        parts.append(f"{_f(v)} * g_RT[{k}]")
    return " + ".join(parts) if parts else "0.0"


def _refc_pow(r: Reaction) -> str:
    d = r.delta_order()
    n = int(round(d))
    if abs(d - n) > 1e-12:
        raise ValueError("non-integer order change")
    if n == 0:
        return ""
    base = "refC" if n > 0 else "refCinv"
    return " * " + " * ".join([base] * abs(n))


def _arr_expr(A: float, b: float, EaR: float) -> str:
    if b == 0.0 and EaR == 0.0:
        return _f(A)
    args = []
    if b != 0.0:
        args.append(f"{_f(b)} * logT")
    if EaR != 0.0:
        args.append(f"{_f(-EaR)} * invT")
    return f"{_f(A)} * exp({' + '.join(args)})"


def _alpha_expr(r: Reaction) -> str:
    s = "mixture" if r.default_eff == 1.0 else f"{_f(r.default_eff)} * mixture"
    for k, e in sorted(r.efficiencies.items()):
        d = e - r.default_eff
        if d != 0.0:
            s += f" + {_f(d)} * sc[{k}]"
    return s


def _troe_fcent(r: Reaction, w: _W, deriv: bool):
    a, T3, T1 = r.troe[0], r.troe[1], r.troe[2]
    T2 = r.troe[3] if len(r.troe) > 3 else None
    f1 = f"{_f(1.0 - a)} * exp(-T / {_f(T3)})" if abs(T3) > 1e-100 and (1.0 - a) != 0 else "0.0"
    f2 = f"{_f(a)} * exp(-T / {_f(T1)})" if abs(T1) > 1e-100 else "0.0"
    f3 = f"exp({_f(-T2)} * invT)" if T2 is not None else "0.0"
    w(f"const double Fc1 = {f1};")
    w(f"const double Fc2 = {f2};")
    w(f"const double Fc3 = {f3};")
    w("const double Fcent = Fc1 + Fc2 + Fc3;")
    w("const double logFcent = log10(Fcent);")
    if deriv:
        d1 = f" - Fc1 / {_f(T3)}" if f1 != "0.0" else ""
        d2 = f" - Fc2 / {_f(T1)}" if f2 != "0.0" else ""
        d3 = f" + Fc3 * {_f(T2)} * invT2" if T2 is not None else ""
        w(f"const double dlogFcentdT = LOG10E / Fcent * (0.0{d1}{d2}{d3});")


def _emit_rates(w: _W, m: Mechanism, with_q: bool):
    """Body shared by production_rate and progress_rates."""
    ns = m.n_species
    w("const double invT = 1.0 / T;")
    w("const double logT = log(T);")
    w("const double refC = 101325.0 / 8.31446 * invT;")
    w("const double refCinv = 1.0 / refC;")
    w("(void)refC; (void)refCinv;")
    w("double mixture = 0.0;")
    w(f"for (int k = 0; k < {ns}; ++k) mixture += sc[k];")
    w(f"double g_RT[{ns}];")
    w("gibbs_RT(g_RT, T);")
    if not with_q:
        w(f"for (int k = 0; k < {ns}; ++k) wdot[k] = 0.0;")
    for j, r in enumerate(m.reactions):
        w(f"{{  // reaction {r.index}: {r.equation}")
        w.ind += 1
        w(f"const double k_f = {_arr_expr(r.rate.A, r.rate.b, r.rate.EaR)};")
        corr = ""
        if r.kind == "three-body":
            w(f"const double Corr = {_alpha_expr(r)};")
            corr = "Corr * "
        elif r.kind == "falloff":
            w(f"const double alpha = {_alpha_expr(r)};")
            w(f"const double redP = alpha / k_f * {_arr_expr(r.low.A, r.low.b, r.low.EaR)};")
            w("const double F = redP / (1.0 + redP);")
            if r.troe is not None:
                w("const double logPred = log10(redP);")
                _troe_fcent(r, w, deriv=False)
                w("const double troe_c = -0.4 - 0.67 * logFcent;")
                w("const double troe_n = 0.75 - 1.27 * logFcent;")
                w("const double troe = (troe_c + logPred) / (troe_n - 0.14 * (troe_c + logPred));")
                w("const double F_troe = exp(LN10 * logFcent / (1.0 + troe * troe));")
                w("const double Corr = F * F_troe;")
            else:
                w("const double Corr = F;")
            corr = "Corr * "
        w(f"const double qf = {corr}k_f * ({_prod_term(r.reactants)});")
        if r.reversible:
            w(f"const double qr = {corr}k_f * exp(-({_kc_arg(r)})){_refc_pow(r)} * ({_prod_term(r.products)});")
            w("const double qdot = qf - qr;")
        else:
            w("const double qdot = qf;")
        if with_q:
            w(f"q[{j}] = qdot;")
        else:
            for k, nu in sorted(r.net_stoich().items()):
                w(f"wdot[{k}] += {_f(nu)} * qdot;" if nu not in (1.0, -1.0)
                  else (f"wdot[{k}] += qdot;" if nu == 1.0 else f"wdot[{k}] -= qdot;"))
        w.ind -= 1
        w("}")


def _emit_jacobian(w: _W, m: Mechanism):
    ns = m.n_species
    n = ns + 1
    w("// Analytic Jacobian in SI-mol units, column-major J[k*(NS+1)+m] = d wdot_m / d c_k,")
    w("// J[NS*(NS+1)+m] = d wdot_m / dT, J[k*(NS+1)+NS] = d Tdot / d c_k, J[last] = d Tdot / dT.")
    w("// CONSP = 1 drops the mixture-concentration derivative of third-body terms and uses")
    w("// cp/h in the temperature row; CONSP = 0 keeps it and uses cv/e.")
    w("template <int CONSP>")
    w("PPC_DEV static inline void jacobian(double* __restrict__ J, const double* __restrict__ sc, const double T) {")
    w.ind += 1
    w(f"for (int i = 0; i < {n * n}; ++i) J[i] = 0.0;")
    w(f"double wdot[{ns}];")
    w(f"for (int k = 0; k < {ns}; ++k) wdot[k] = 0.0;")
    w("const double invT = 1.0 / T;")
    w("const double invT2 = invT * invT;")
    w("const double logT = log(T);")
    w("const double refC = 101325.0 / 8.31446 * invT;")
    w("const double refCinv = 1.0 / refC;")
    w("(void)refC; (void)refCinv; (void)invT2;")
    w("double mixture = 0.0;")
    w(f"for (int k = 0; k < {ns}; ++k) mixture += sc[k];")
    w(f"double g_RT[{ns}], hRT[{ns}];")
    w("gibbs_RT(g_RT, T);")
    w("h_RT(hRT, T);")
    for r in m.reactions:
        w(f"{{  // reaction {r.index}: {r.equation}")
        w.ind += 1
        w(f"const double phi_f = {_prod_term(r.reactants)};")
        w(f"double k_f = {_arr_expr(r.rate.A, r.rate.b, r.rate.EaR)};")
        w(f"const double dlnkfdT = {_f(r.rate.b)} * invT + {_f(r.rate.EaR)} * invT2;")
        if r.kind == "falloff":
            w(f"const double alpha = {_alpha_expr(r)};")
            w(f"const double k_0 = {_arr_expr(r.low.A, r.low.b, r.low.EaR)};")
            w("const double Pr = alpha / k_f * k_0;")
            w("const double fPr = Pr / (1.0 + Pr);")
            w(f"const double dlnk0dT = {_f(r.low.b)} * invT + {_f(r.low.EaR)} * invT2;")
            w("const double dlogPrdT = LOG10E * (dlnk0dT - dlnkfdT);")
            w("const double dlogfPrdT = dlogPrdT / (1.0 + Pr);")
            if r.troe is not None:
                w("const double logPr = log10(Pr);")
                _troe_fcent(r, w, deriv=True)
                w("const double troe_c = -0.4 - 0.67 * logFcent;")
                w("const double troe_n = 0.75 - 1.27 * logFcent;")
                w("const double troePr_den = 1.0 / (troe_n - 0.14 * (troe_c + logPr));")
                w("const double troePr = (troe_c + logPr) * troePr_den;")
                w("const double troe = 1.0 / (1.0 + troePr * troePr);")
                w("const double F = exp(LN10 * logFcent * troe);")
                w("const double dlogFdcn_fac = 2.0 * logFcent * troe * troe * troePr * troePr_den;")
                w("const double dlogFdc = -troe_n * dlogFdcn_fac * troePr_den;")
                w("const double dlogFdn = dlogFdcn_fac * troePr;")
                w("const double dlogFdlogPr = dlogFdc;")
                w("const double dlogFdT = dlogFcentdT * (troe - 0.67 * dlogFdc - 1.27 * dlogFdn) + dlogFdlogPr * dlogPrdT;")
            else:
                w("const double F = 1.0;")
                w("const double dlogFdlogPr = 0.0;")
                w("const double dlogFdT = 0.0;")
        elif r.kind == "three-body":
            w(f"const double alpha = {_alpha_expr(r)};")
        if r.reversible:
            w(f"const double phi_r = {_prod_term(r.products)};")
            # Kc = exp(sum nu' g - sum nu'' g) * refC^(-delta) ; k_r = k_f / Kc
            w(f"double k_r = k_f * exp(-({_kc_arg(r)})){_refc_pow(r)};")
            # d ln Kc / dT = (sum nu'' h - sum nu' h)/T - delta_order... sign as reference
            hr = " + ".join(f"{_f(v)} * hRT[{k}]" for k, v in sorted(r.products.items()))
            hf = " + ".join(f"{_f(v)} * hRT[{k}]" for k, v in sorted(r.reactants.items()))
            d = r.delta_order()
            w(f"const double dlnKcdT = invT * (-({hf}) + ({hr}) + {_f(d)});")
            w("const double dkrdT = (dlnkfdT - dlnKcdT) * k_r;")
        else:
            w("const double phi_r = 0.0;")
            w("double k_r = 0.0;")
            w("const double dkrdT = 0.0;")
        if r.kind == "elementary":
            w("const double q = k_f * phi_f - k_r * phi_r;")
            w("const double dqdT = dlnkfdT * k_f * phi_f - dkrdT * phi_r;")
            fac = None
        elif r.kind == "three-body":
            w("const double q_nocor = k_f * phi_f - k_r * phi_r;")
            w("const double q = alpha * q_nocor;")
            w("const double dqdT = alpha * (dlnkfdT * k_f * phi_f - dkrdT * phi_r);")
            w("k_f *= alpha; k_r *= alpha;")
            fac = "q_nocor"
        else:
            w("const double q_nocor = k_f * phi_f - k_r * phi_r;")
            w("const double Corr = fPr * F;")
            w("const double q = Corr * q_nocor;")
            w("const double dlnCorrdT = LN10 * (dlogfPrdT + dlogFdT);")
            w("const double dqdT = Corr * (dlnkfdT * k_f * phi_f - dkrdT * phi_r) + dlnCorrdT * q;")
            w("k_f *= Corr; k_r *= Corr;")
            w("const double dcdc_fac = q / alpha * (1.0 / (Pr + 1.0) + dlogFdlogPr);")
            fac = "dcdc_fac"
        w("(void)phi_r; (void)k_r;")
        net = sorted(r.net_stoich().items())
        for k, nu in net:
            w(f"wdot[{k}] += {_f(nu)} * q;")
        # species derivative contributions
        involved = sorted(set(r.reactants) | (set(r.products) if r.reversible else set()))

        def dq_phi(k: int) -> str:
            parts = []
            if k in r.reactants:
                parts.append(f"k_f * {_dphi(r.reactants, k)}")
            if r.reversible and k in r.products:
                parts.append(f"- k_r * {_dphi(r.products, k)}")
            return " ".join(("+ " + p if not p.startswith("-") else p) for p in parts)

        def emit_col(k: int, expr: str):
            w(f"{{ const double dqdci = {expr};")
            w.ind += 1
            for mm, nu in net:
                w(f"J[{k * n + mm}] += {_f(nu)} * dqdci;")
            w.ind -= 1
            w("}")

        if fac is None:
            for k in involved:
                emit_col(k, "0.0 " + dq_phi(k))
        else:
            # constant-pressure branch: only species with efficiency != 1 (as eff-1) and
            # the species of the rate expression
            effk = {k: e - r.default_eff for k, e in r.efficiencies.items() if e != r.default_eff}
            w("if (CONSP == 1) {")
            w.ind += 1
            for k in sorted(set(effk) | set(involved)):
                e = f"{_f(effk[k])} * {fac} " if k in effk else "0.0 "
                emit_col(k, e + (dq_phi(k) if k in involved else ""))
            w.ind -= 1
            w("} else {")
            w.ind += 1
            w(f"for (int k = 0; k < {ns}; ++k) {{")
            w.ind += 1
            w(f"const double dqdci = {_f(r.default_eff)} * {fac};")
            for mm, nu in net:
                w(f"J[k * {n} + {mm}] += {_f(nu)} * dqdci;")
            w.ind -= 1
            w("}")
            for k in sorted(set(effk) | set(involved)):
                e = f"{_f(effk[k])} * {fac} " if k in effk else "0.0 "
                emit_col(k, e + (dq_phi(k) if k in involved else ""))
            w.ind -= 1
            w("}")
        for mm, nu in net:
            w(f"J[{ns * n + mm}] += {_f(nu)} * dqdT;")
        w.ind -= 1
        w("}")
    # temperature row
    w(f"double c_R[{ns}], dcRdT[{ns}];")
    w("if (CONSP == 1) { cp_R(c_R, T); } else { cv_R(c_R, T); }")
    w("dcpR_dT(dcRdT, T);")
    w(f"double ehRT[{ns}];")
    w(f"for (int k = 0; k < {ns}; ++k) ehRT[k] = (CONSP == 1) ? hRT[k] : hRT[k] - 1.0;")
    w("double cmix = 0.0, ehmix = 0.0, dcmixdT = 0.0, dehmixdT = 0.0;")
    w(f"for (int k = 0; k < {ns}; ++k) {{")
    w("  cmix += c_R[k] * sc[k];")
    w("  dcmixdT += dcRdT[k] * sc[k];")
    w("  ehmix += ehRT[k] * wdot[k];")
    w(f"  dehmixdT += invT * (c_R[k] - ehRT[k]) * wdot[k] + ehRT[k] * J[{ns * n} + k];")
    w("}")
    w("const double cmixinv = 1.0 / cmix;")
    w("const double tmp1 = ehmix * cmixinv;")
    w("const double tmp3 = cmixinv * T;")
    w("const double tmp2 = tmp1 * tmp3;")
    w(f"for (int k = 0; k < {ns}; ++k) {{")
    w("  double dehmixdc = 0.0;")
    w(f"  for (int mm = 0; mm < {ns}; ++mm) dehmixdc += ehRT[mm] * J[k * {n} + mm];")
    w(f"  J[k * {n} + {ns}] = tmp2 * c_R[k] - tmp3 * dehmixdc;")
    w("}")
    w(f"J[{n * n - 1}] = -tmp1 + tmp2 * dcmixdT - tmp3 * dehmixdT;")
    w.ind -= 1
    w("}")


def emit(m: Mechanism, source: str) -> str:
    ns, nr = m.n_species, m.n_reactions
    w = _W()
    w(f"// Generated by pelephysics_amd.mechgen.emit_hip from {source}. Do not edit.")
    w(f"// {m.name}: {ns} species, {nr} reactions.")
    w("#pragma once")
    w('#include "../ppc_common.h"')
    w()
    w("namespace ppc {")
    w("namespace mech {")
    w()
    w(f"struct {m.name} {{")
    w.ind += 1
    w(f"static constexpr int NS = {ns};")
    w(f"static constexpr int NR = {nr};")
    w(f"static constexpr int NEQ = {ns + 1};")
    w()
    w("// molecular weights (g/mol) and their inverses")
    w("PPC_HD static constexpr double mw(int k) {")
    w(f"  constexpr double t[{ns}] = {{{', '.join(_f(s.mw) for s in m.species)}}};")
    w("  return t[k];")
    w("}")
    w("PPC_HD static constexpr double imw(int k) {")
    w(f"  constexpr double t[{ns}] = {{{', '.join(_f(s.imw) for s in m.species)}}};")
    w("  return t[k];")
    w("}")
    w("// runtime-indexed variants (constant memory)")
    w("PPC_DEV static double mw_rt(int k);")
    w("PPC_DEV static double imw_rt(int k);")
    w()
    _emit_species_fn(w, m, "cp_R", _cp_expr, "species cp/R (NASA7)")
    _emit_species_fn(w, m, "cv_R", _cv_expr, "species cv/R")
    _emit_species_fn(w, m, "h_RT", _h_expr, "species h/(RT)")
    _emit_species_fn(w, m, "e_RT", _e_expr, "species e/(RT)")
    _emit_species_fn(w, m, "gibbs_RT", _g_expr, "species g/(RT)", needs_log=True)
    _emit_species_fn(w, m, "dcpR_dT", _dcp_expr, "species d(cp/R)/dT = d(cv/R)/dT")

    # fused mixture functions (mass units)
    for nm, hfn, cfn, doc in (("h_cp_mass", _h_expr, _cp_expr, "mixture enthalpy (erg/g) and cp (erg/g/K)"),
                              ("e_cv_mass", _e_expr, _cv_expr, "mixture internal energy (erg/g) and cv (erg/g/K)")):
        w(f"// {doc} from mass fractions")
        w(f"PPC_DEV static inline void {nm}(const double T, const double* __restrict__ Y, double& hm, double& cm) {{")
        w.ind += 1
        w("const double invT = 1.0 / T;")
        w("(void)invT;")
        w("double hs = 0.0, cs = 0.0;")
        for tmid, ks in _thermo_groups(m).items():
            w(f"{{ const bool lo = T < {_f(tmid)};")
            w.ind += 1
            for k in ks:
                s = m.species[k]
                w(f"{{ const double yw = Y[{k}] * {_f(s.imw)};")
                w(f"  hs += yw * ({hfn(s)});")
                w(f"  cs += yw * ({cfn(s)}); }}")
            w.ind -= 1
            w("}")
        w("hm = hs * (RU * T);")
        w("cm = cs * RU;")
        w.ind -= 1
        w("}")
        w()

    w("// molar production rates wdot (mol/m^3/s) from concentrations sc (mol/m^3)")
    w("PPC_DEV static inline void production_rate(double* __restrict__ wdot, const double* __restrict__ sc, const double T) {")
    w.ind += 1
    _emit_rates(w, m, with_q=False)
    w.ind -= 1
    w("}")
    w()
    w("// net rate of progress q_j (mol/m^3/s) of every reaction")
    w("PPC_DEV static inline void progress_rates(double* __restrict__ q, const double* __restrict__ sc, const double T) {")
    w.ind += 1
    _emit_rates(w, m, with_q=True)
    w.ind -= 1
    w("}")
    w()
    w("// net stoichiometric coefficients of reaction j acting on species k, as a sparse list")
    nnz = []
    for j, r in enumerate(m.reactions):
        for k, nu in sorted(r.net_stoich().items()):
            nnz.append((j, k, nu))
    w(f"static constexpr int NNZ_STOICH = {len(nnz)};")
    w("PPC_DEV static inline void stoich_entry(int i, int& j, int& k, double& nu) {")
    w(f"  constexpr int tj[{len(nnz)}] = {{{', '.join(str(a) for a, _, _ in nnz)}}};")
    w(f"  constexpr int tk[{len(nnz)}] = {{{', '.join(str(b) for _, b, _ in nnz)}}};")
    w(f"  constexpr double tn[{len(nnz)}] = {{{', '.join(_f(c) for _, _, c in nnz)}}};")
    w("  j = tj[i]; k = tk[i]; nu = tn[i];")
    w("}")
    w()
    _emit_jacobian(w, m)
    w.ind -= 1
    w("};")
    w()
    w(f"namespace detail_{m.name} {{")
    w(f"static __constant__ double kMW[{ns}] = {{{', '.join(_f(s.mw) for s in m.species)}}};")
    w(f"static __constant__ double kIMW[{ns}] = {{{', '.join(_f(s.imw) for s in m.species)}}};")
    w("}")
    w(f"PPC_DEV double {m.name}::mw_rt(int k) {{ return detail_{m.name}::kMW[k]; }}")
    w(f"PPC_DEV double {m.name}::imw_rt(int k) {{ return detail_{m.name}::kIMW[k]; }}")
    w()
    w("}  // namespace mech")
    w("}  // namespace ppc")
    return w.text()
