"""Oracle kinetics pinned against the reference literals and against invariants.

* per reaction, at the fixture temperatures: forward rate = k_f * prod c^nu'
  (elementary), k_f * alpha * prod c^nu' (three-body, alpha from the literal
  enhancement), and at the Troe midpoint log Pr = -c the factor F = F_cent
  (falloff);
* every falloff reaction at five reduced pressures around the Troe midpoint
  (log Pr = -c - 2 .. -c + 2), where the troe_n and 0.14 terms of the
  broadening factor matter, against the factor the reference's own statement
  chain gives (the ``falloff`` entries of the fixture);
* detailed balance q = 0 when prod c''^nu'' / prod c'^nu' equals
  K_c formed from the literal g/RT and the reference's reference
  concentration 101325 / 8.31446 / T;
* element and mass conservation of the production rates;
* the Jacobian known-answer test of Testing/Exec/Jacobian/main.cpp
  (srand(42) concentrations, T = 1000 K, central differences with step 1e-4,
  sum(err^2)/sum(J_fd^2) <= 1e-12), extended to the temperature column and
  row; the constant-pressure variant differs from the exact derivative by a
  row-independent term (the dropped d(mixture)/dc).
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
MECHS = ["LiDryer", "drm19", "dodecane_lu", "heptane_lu_88sk"]


def _fix(mech):
    with open(os.path.join(HERE, "golden", f"{mech}_literals.json")) as f:
        return json.load(f)


def _conc(rng, ns):
    return rng.uniform(0.5, 2.0, ns)


def _kc(R, d, ti):
    T = d["tgrid"][ti]
    g = np.array(d["thermo"]["gibbs"][ti])
    refC = 101325 / 8.31446 / T
    dg = sum(v * g[k] for k, v in R.reactants.items()) - sum(v * g[k] for k, v in R.products.items())
    return math.exp(dg) * refC ** (-R.delta_order())


@pytest.mark.parametrize("mech", MECHS)
def test_net_rates_against_literals(models, oracles, mech):
    """q_j from the literal k_f, k_0, F_cent, enhancement and g/RT of the reference."""
    m, O, d = models[mech], oracles[mech], _fix(mech)
    rng = np.random.default_rng(3)
    for r in d["reactions"]:
        j = r["index"]
        R = m.reactions[j]
        for ti in range(0, len(d["tgrid"]), 3):
            T = d["tgrid"][ti]
            c = _conc(rng, m.n_species)
            kf = r["kf"][ti]
            corr = 1.0
            if r["enhancement"] is not None:
                alpha = r["mixture_coef"] * c.sum() + sum(v * c[int(k)] for k, v in r["enhancement"].items())
                if r["k0"] is None:
                    corr = alpha
                else:
                    # place Pr at the Troe midpoint log10 Pr = -troe_c, where F = F_cent
                    if r["Fcent"] is not None:
                        fc = r["Fcent"][ti]
                        logpr = 0.4 + 0.67 * math.log10(fc)
                    else:
                        fc, logpr = 1.0, 0.0
                    c *= (10 ** logpr * kf / r["k0"][ti]) / alpha
                    Pr = 10 ** logpr
                    corr = Pr / (1 + Pr) * fc
            phi_f = np.prod([c[int(k)] ** v for k, v in r["reactants"].items()])
            expect = corr * kf * phi_f
            if R.reversible:
                phi_r = np.prod([c[k] ** v for k, v in R.products.items()])
                expect -= corr * kf / _kc(R, d, ti) * phi_r
            q = O.qdot(c, T)[j]
            # the literal g/RT carry ~1e-9 relative rounding: K_c to ~1e-7
            assert q == pytest.approx(expect, rel=1e-6, abs=1e-10 * corr * kf * phi_f), (R.equation, T)


def falloff_cases(m, d, rng, tis=range(8)):
    """(j, T, c, expected q) at the literal falloff points of every falloff
    reaction: c scaled so that the reduced pressure is the fixture's Pr."""
    out = []
    for r in d["reactions"]:
        if r["falloff"] is None:
            continue
        j = r["index"]
        R = m.reactions[j]
        for ti in tis:
            T = d["tgrid"][ti]
            kf, k0 = r["kf"][ti], r["k0"][ti]
            for logpr, factor in r["falloff"][ti]:
                c = _conc(rng, m.n_species)
                alpha = r["mixture_coef"] * c.sum() + sum(v * c[int(k)] for k, v in r["enhancement"].items())
                c *= (10 ** logpr * kf / k0) / alpha
                expect = factor * kf * np.prod([c[int(k)] ** v for k, v in r["reactants"].items()])
                if R.reversible:
                    expect -= factor * kf / _kc(R, d, ti) * np.prod([c[k] ** v for k, v in R.products.items()])
                out.append((j, T, c, expect, factor * kf))
    return out


@pytest.mark.parametrize("mech", MECHS)
def test_falloff_off_midpoint_against_literals(models, oracles, mech):
    """Troe and Lindemann factors away from the Troe midpoint, from the
    reference's literal statement chain (mechanism.H, e.g. drm19 :4446-4461)."""
    m, O, d = models[mech], oracles[mech], _fix(mech)
    cases = falloff_cases(m, d, np.random.default_rng(6))
    assert len(cases) == 40 * sum(r["falloff"] is not None for r in d["reactions"])
    for j, T, c, expect, scale in cases:
        phi = np.prod([c[k] ** v for k, v in m.reactions[j].reactants.items()])
        q = O.qdot(c, T)[j]
        assert q == pytest.approx(expect, rel=1e-6, abs=1e-9 * scale * phi), (m.reactions[j].equation, T)


def test_falloff_fixture_moves_off_midpoint():
    """The off-midpoint points change the Troe factor by more than the test
    tolerance when troe_n or the 0.14 coefficient changes (what a dropped or
    wrong term would do): recompute one drm19 block's chain with n + 0.1."""
    d = _fix("drm19")
    r = next(x for x in d["reactions"] if x["Fcent"] is not None)
    ti = 4
    logfc = math.log10(r["Fcent"][ti])
    c = -0.4 - 0.67 * logfc
    for (logpr, factor), off in zip(r["falloff"][ti], (-2, -1, 0, 1, 2)):
        Pr = 10 ** logpr
        for dn, d14 in ((0.0, 0.0), (0.1, 0.0), (0.0, 0.02)):
            n = 0.75 - 1.27 * logfc + dn
            troe = (c + logpr) / (n - (0.14 + d14) * (c + logpr))
            f = Pr / (1 + Pr) * 10 ** (logfc / (1 + troe * troe))
            if dn == 0.0 and d14 == 0.0:
                assert f == pytest.approx(factor, rel=1e-12)
            elif off != 0:
                assert abs(f / factor - 1) > 1e-5


@pytest.mark.parametrize("mech", MECHS)
def test_detailed_balance_against_literal_gibbs(models, oracles, mech):
    m, O, d = models[mech], oracles[mech], _fix(mech)
    rng = np.random.default_rng(4)
    for r in d["reactions"]:
        j = r["index"]
        R = m.reactions[j]
        if not R.reversible:
            continue
        for ti in (1, 4, 6):
            T = d["tgrid"][ti]
            Kc = _kc(R, d, ti)
            c = _conc(rng, m.n_species)
            # solve for one species so that prod c''^nu'' / prod c'^nu' = Kc
            nu = R.net_stoich()
            k0 = max(nu, key=lambda k: abs(nu[k]))
            c[k0] = 1.0
            ratio = np.prod([c[k] ** v for k, v in R.products.items()]) / \
                np.prod([c[k] ** v for k, v in R.reactants.items()])
            c[k0] = (Kc / ratio) ** (1.0 / nu[k0])
            q = O.qdot(c, T)[j]
            phi_f = np.prod([c[k] ** v for k, v in R.reactants.items()])
            qf = r["kf"][ti] * phi_f * (1.0 if R.kind == "elementary" else 1e3)
            assert abs(q) <= 1e-6 * abs(qf) + 1e-300, (R.equation, T, q, qf)


@pytest.mark.parametrize("mech", MECHS)
def test_element_and_mass_conservation(models, oracles, mech):
    m, O = models[mech], oracles[mech]
    rng = np.random.default_rng(5)
    E = np.array([[s.composition.get(e, 0.0) for e in m.elements] for s in m.species])
    mw = np.array([s.mw for s in m.species])
    for T in (400.0, 1100.0, 2300.0):
        c = rng.uniform(0.0, 5.0, m.n_species)
        w = O.wdot(c, T)
        scale = np.abs(w).max()
        assert np.abs(E.T @ w).max() <= 1e-12 * scale * E.max()
        assert abs(mw @ w) <= 1e-12 * scale * mw.max()


def _glibc_srand42(n):
    """The reference KAT's concentrations: srand(42); rand()/RAND_MAX + 1e-4."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(42)
    RAND_MAX = 2147483647
    return np.array([libc.rand() / RAND_MAX + 1e-4 for _ in range(n)])


def _fd_jacobian(O, sc, T, pert=1e-4):
    ns = sc.size
    J = np.zeros((ns + 1, ns))  # J[k, m] = d wdot_m / d c_k (k = ns: T)
    for k in range(ns + 1):
        if k < ns:
            a, b = sc.copy(), sc.copy()
            a[k] += pert
            b[k] -= pert
            J[k] = (O.wdot(a, T) - O.wdot(b, T)) / (2 * pert)
        else:
            h = 1e-3
            J[k] = (O.wdot(sc, T + h) - O.wdot(sc, T - h)) / (2 * h)
    return J


def _tdot(O, sc, T, consP):
    """Chemistry-only temperature rate -T sum(eh_k w_k) / sum(c_k c_R,k) (molar form)."""
    w = O.wdot(sc, T)
    hrt = O.h_RT(T)
    cp = O.cp_R(T)
    eh = hrt if consP else hrt - 1.0
    cr = cp if consP else cp - 1.0
    return -T * (eh @ w) / (cr @ sc)


@pytest.mark.parametrize("mech", MECHS)
def test_jacobian_kat_exact_derivative(models, oracles, mech):
    """consP = 0: exact derivative; species block, T column and T row against central FD."""
    m, O = models[mech], oracles[mech]
    ns, n = m.n_species, m.n_species + 1
    sc = _glibc_srand42(ns)
    T = 1000.0
    J = O.jacobian(sc, T, 0).reshape(n, n)  # J[k, m] (column-major storage)
    Jfd = _fd_jacobian(O, sc, T)
    err = J[:ns, :ns] - Jfd[:ns]
    assert (err ** 2).sum() / (Jfd[:ns] ** 2).sum() <= 1e-12
    # temperature column and row away from the NASA7 midpoint (1000 K), where
    # a central difference in T would straddle the two polynomial fits
    T = 1234.5
    J = O.jacobian(sc, T, 0).reshape(n, n)
    Jfd = _fd_jacobian(O, sc, T)
    np.testing.assert_allclose(J[ns, :ns], Jfd[ns], rtol=1e-6, atol=1e-9 * np.abs(Jfd[ns]).max())
    # temperature row: d Tdot / d c_k and d Tdot / dT (e-form)
    row = np.zeros(n)
    for k in range(n):
        if k < ns:
            p = 1e-4
            a, b = sc.copy(), sc.copy()
            a[k] += p
            b[k] -= p
            row[k] = (_tdot(O, a, T, 0) - _tdot(O, b, T, 0)) / (2 * p)
        else:
            h = 1e-3
            row[k] = (_tdot(O, sc, T + h, 0) - _tdot(O, sc, T - h, 0)) / (2 * h)
    np.testing.assert_allclose(J[:, ns], row, rtol=1e-6, atol=1e-9 * np.abs(row).max())


@pytest.mark.parametrize("mech", MECHS)
def test_jacobian_constant_pressure_convention(models, oracles, mech):
    """consP = 1 drops d(mixture)/dc of third-body terms: J_fd - J_consP is the same for every c_k."""
    m, O = models[mech], oracles[mech]
    ns, n = m.n_species, m.n_species + 1
    sc = _glibc_srand42(ns)
    T = 1234.5
    Jp = O.jacobian(sc, T, 1).reshape(n, n)
    Jfd = _fd_jacobian(O, sc, T)
    diff = Jfd[:ns] - Jp[:ns, :ns]
    scale = np.abs(Jfd).max()
    assert np.abs(diff - diff[0][None, :]).max() <= 1e-8 * scale
    assert np.abs(diff[0]).max() > 1e-6 * scale  # the dropped term is not empty
    # the dT column does not depend on the convention
    np.testing.assert_allclose(Jp[ns, :ns], Jfd[ns], rtol=1e-6, atol=1e-9 * np.abs(Jfd[ns]).max())
    # the T row uses cp/h and the consP species block
    cp, hrt, w = O.cp_R(T), O.h_RT(T), O.wdot(sc, T)
    cmix = cp @ sc
    tmp2 = T * (hrt @ w) / cmix ** 2
    row = tmp2 * cp - T / cmix * (Jp[:ns, :ns] @ hrt)
    np.testing.assert_allclose(Jp[:ns, ns], row, rtol=1e-10, atol=1e-12 * np.abs(row).max())
