"""The YAML-derived mechanism tables against the literals of the reference headers.

Fixtures: tests/golden/<mech>_literals.json, written by
tests/golden/extract_mechanism_literals.py from Mechanisms/<mech>/mechanism.H
(rate constants, efficiencies, stoichiometry, Troe F_cent, thermo functions
evaluated on a temperature grid).
"""
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


def _arr(r, T):
    return r.A * math.exp(r.b * math.log(T) - r.EaR / T)


@pytest.mark.parametrize("mech", MECHS)
def test_molecular_weights(models, mech):
    d = _fix(mech)
    assert [s.mw for s in models[mech].species] == d["mw"]
    for s in models[mech].species:
        assert abs(s.imw - 1.0 / s.mw) < 1e-15


@pytest.mark.parametrize("mech", MECHS)
def test_rate_constants_and_structure(models, mech):
    m = models[mech]
    d = _fix(mech)
    assert len(d["reactions"]) == m.n_reactions
    for r in d["reactions"]:
        R = m.reactions[r["index"]]
        for T, kf in zip(d["tgrid"], r["kf"]):
            assert _arr(R.rate, T) == pytest.approx(kf, rel=1e-12), R.equation
        if r["k0"] is not None:
            assert R.kind == "falloff"
            for T, k0 in zip(d["tgrid"], r["k0"]):
                assert _arr(R.low, T) == pytest.approx(k0, rel=1e-12), R.equation
        else:
            assert R.kind != "falloff"
        if r["Fcent"] is not None:
            a, T3, T1 = R.troe[:3]
            for T, fc in zip(d["tgrid"], r["Fcent"]):
                mine = (1 - a) * math.exp(-T / T3) + a * math.exp(-T / T1)
                if len(R.troe) == 4:
                    mine += math.exp(-R.troe[3] / T)
                assert mine == pytest.approx(fc, rel=1e-12, abs=1e-300), R.equation
        if r["enhancement"] is not None:
            assert R.has_third_body
            assert r["mixture_coef"] == R.default_eff
            enh = {int(k): v for k, v in r["enhancement"].items()}
            mine = {k: e - R.default_eff for k, e in R.efficiencies.items() if e != R.default_eff}
            assert enh.keys() == mine.keys(), R.equation
            for k in enh:
                assert enh[k] == pytest.approx(mine[k], abs=1e-12)
        else:
            assert not R.has_third_body
        assert {int(k): v for k, v in r["reactants"].items()} == R.reactants
        net = {int(k): v for k, v in r["net"].items() if v != 0.0}
        assert net == R.net_stoich()
        assert r["reversible"] == R.reversible
        if R.reversible:
            assert r["refc_power"] == pytest.approx(R.delta_order())


@pytest.mark.parametrize("mech", MECHS)
def test_thermo_functions(oracles, mech):
    """Oracle cp/R, h/RT, g/RT against the reference's thermo functions on the grid.

    The reference prints the h/RT coefficients a_k/k with 9 significant digits,
    hence the 1e-7 relative tolerance on h/RT (absolute 1e-6 near h = 0); its
    g/RT constants are differences of the YAML coefficients rounded to 16
    digits (1e-8 relative)."""
    O = oracles[mech]
    d = _fix(mech)
    for i, T in enumerate(d["tgrid"]):
        np.testing.assert_allclose(O.cp_R(T), d["thermo"]["cp_R"][i], rtol=1e-9)
        np.testing.assert_allclose(O.h_RT(T), d["thermo"]["speciesEnthalpy"][i], rtol=1e-7, atol=1e-6)
        np.testing.assert_allclose(O.g_RT(T), d["thermo"]["gibbs"][i], rtol=1e-8, atol=1e-7)


def test_json_matches_yaml_when_reference_present(models):
    src = "/root/reference/Mechanisms"
    if not os.path.isdir(src):
        pytest.skip("reference checkout not present")
    from pelephysics_amd.mechgen.model import load_yaml
    for mech in MECHS:
        m = load_yaml(os.path.join(src, mech, "mechanism.yaml"), mech)
        assert m.to_json() == models[mech].to_json()
