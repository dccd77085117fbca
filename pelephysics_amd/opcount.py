"""Static FP64 operation counts of the per-cell chemistry functions.

Counts follow the emitted device code (pelephysics_amd/mechgen/emit_hip.py):
add/sub/mul = 1, fused multiply-add = 2, divide = 1, and each transcendental
(exp, log, log10, pow, sqrt) = 1 (transcendentals are also reported
separately).  They are algorithmic counts of the formulas, not instruction
counts of the compiled code.

Per cell, one integration performs n_rhs right-hand sides, n_jac Jacobians,
n_lu LU factorisations and n_newton Newton solves (each a pair of
triangular solves); the kernels report these counters.  Then

    W_cell = n_rhs * F_rhs + n_jac * F_jac + n_lu * (2/3) n^3 + n_newton * 2 n^2

with n = N + 1.  The T inversion inside F_rhs is counted with the 2 bound
evaluations plus ``TINV_ITERS`` Newton iterations: the mean number of energy
evaluations the Newton loop takes inside the integrators, counted on the
oracle over a sample of the drm19 64^3 grid (dt 1e-5 s, rtol 1e-10): 1.003
on the first call and 1.008 on a chained call, since the state carries T to
well within the loop's 1e-6 K tolerance (tests/test_opcount.py keeps the
figure within 0.5 of the count).
"""
from __future__ import annotations

from dataclasses import dataclass

from .mechgen.model import Mechanism

TINV_ITERS = 1  # energy evaluations of the T inversion's Newton loop inside the RHS (counted)


@dataclass
class OpCounts:
    n: int
    thermo_species: int  # one species polynomial (cp/R or h/RT)
    mixture_h_cp: int  # fused mixture h and cp (mass units)
    gibbs: int
    rates: int  # production_rate
    rates_transc: int
    jac: int
    jac_transc: int
    rhs: int  # full f_kernel_spec
    rhs_transc: int

    @property
    def lu(self) -> float:
        return 2.0 / 3.0 * self.n ** 3

    @property
    def solve(self) -> float:
        return 2.0 * self.n ** 2


def counts(m: Mechanism) -> OpCounts:
    ns, n = m.n_species, m.n_species + 1
    poly = 8  # Horner degree-4 polynomial: 4 FMA
    h_cp_species = 2 * poly + 2 + 4  # h and cp polynomials, 1/T term, weighting (2 FMA)
    mixture = ns * h_cp_species + 4
    gibbs = ns * (2 * 3 + 6 + 2) + 2  # degree-3 Horner, a6/T, a1 lnT, constant; + log
    rates = 3 + ns + gibbs + 1  # invT, logT, refC, mixture, gibbs (+log)
    rt = 1 + 1 + 1  # log, 1/T, refC
    for r in m.reactions:
        rates += 4  # k_f exponent
        rt += 1 if (r.rate.b != 0.0 or r.rate.EaR != 0.0) else 0
        if r.has_third_body:
            rates += 2 * sum(1 for e in r.efficiencies.values() if e != r.default_eff)
        if r.kind == "falloff":
            rates += 4 + 3 + 2  # k0 exponent, redP, F
            rt += 1
            if r.troe is not None:
                rates += 20
                rt += 4 + 2  # 2-3 exp for Fcent, log10 Fcent, log10 Pr, exp
        nr = sum(int(v) for v in r.reactants.values())
        npd = sum(int(v) for v in r.products.values())
        rates += nr + 1  # qf
        if r.reversible:
            rates += 2 * len(r.net_stoich()) + npd + 3
            rt += 1
        rates += 2 * len(r.net_stoich())  # wdot accumulation
    # f_kernel_spec: rho, Y, T inversion (2 bounds + TINV_ITERS iterations), e_k/h_k, cp,
    # concentrations, rates, mass rates, T equation
    rhs = 2 * ns + 3 + (2 + TINV_ITERS + 1) * mixture + ns * (poly + 4) + 3 * ns + rates + 2 * ns + 4 * ns + 3
    rhs_t = rt
    # Jacobian: rates work + derivative terms + T row
    jac = rates + gibbs
    jt = rt
    for r in m.reactions:
        jac += 10
        involved = len(r.reactants) + (len(r.products) if r.reversible else 0)
        if r.has_third_body:
            involved += sum(1 for e in r.efficiencies.values() if e != r.default_eff)
        if r.kind == "falloff":
            jac += 40
            if r.troe is not None:
                jt += 1
        jac += involved * (3 + 2 * len(r.net_stoich()))
        jac += 2 * len(r.net_stoich())
    jac += ns * (2 * poly) + ns * 10 + ns * ns * 2 + ns * 4 + 10
    return OpCounts(n, h_cp_species, mixture, gibbs, rates, rt, jac, jt, rhs, rhs_t)


def work_per_cell(oc: OpCounts, nrhs: float, njac: float, nlu: float, nnewton: float) -> dict:
    """Algorithmic FP64 and FP32 flops of one cell (LU counted in the precision it runs in)."""
    rhs = nrhs * oc.rhs
    jac = njac * oc.jac
    lu = nlu * oc.lu
    sol = nnewton * oc.solve
    return {"rhs": rhs, "jac": jac, "lu": lu, "solve": sol, "total": rhs + jac + lu + sol}
