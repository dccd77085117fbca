"""GPU parity of the mechanism functions the integrators call, through the C-ABI:
productionRate, the net rates of progress and aJacobian (both consP
conventions) of every compiled implementation (generated straight-line code
and constant-table walks) against the oracle, and the falloff reactions
against the reference's own literal statement chains.

* falloff: at five reduced pressures around the Troe midpoint per grid
  temperature (log Pr = -c - 2 .. -c + 2, where the troe_n and 0.14 terms of
  the broadening factor matter), the GPU's q_j equals the value of the
  reference's statements (fixtures tests/golden/<mech>_literals.json, written
  by extract_mechanism_literals.py) to 1e-6 relative, the tolerance of the
  literal g/RT; and equals the oracle's q_j to 1e-12;
* random states over 300-3000 K with concentrations spanning six decades:
  wdot and every Jacobian entry against the oracle, normwise per cell to
  1e-11 (device and host exp/log differ by a few ulp).
"""
import numpy as np
import pytest

from test_oracle_kinetics import falloff_cases, _fix

pytestmark = pytest.mark.gpu

IMPLS = [("LiDryer", "LiDryer"), ("LiDryer", "LiDryer_table"), ("drm19", "drm19"), ("drm19", "drm19_table"),
         ("dodecane_lu", "dodecane_lu"), ("heptane_lu_88sk", "heptane_lu_88sk")]


def _handle(kernels):
    import torch
    from pelephysics_amd.reactor import ReactorBase
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    R = ReactorBase.create("ReactorBDF")
    R.init(2, 1, mechanism=kernels)
    return R


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda:0")


@pytest.mark.parametrize("mech,kernels", IMPLS)
def test_falloff_rates_against_literals(models, oracles, mech, kernels):
    m, O, d = models[mech], oracles[mech], _fix(mech)
    cases = falloff_cases(m, d, np.random.default_rng(8))
    R = _handle(kernels)
    sc = np.array([c for _, _, c, _, _ in cases])
    T = np.array([t for _, t, _, _, _ in cases])
    q = R.progress_rate(_dev(sc), _dev(T)).cpu().numpy()
    for i, (j, t, c, expect, scale) in enumerate(cases):
        phi = np.prod([c[k] ** v for k, v in m.reactions[j].reactants.items()])
        assert q[i, j] == pytest.approx(expect, rel=1e-6, abs=1e-9 * scale * phi), (m.reactions[j].equation, t)
        qo = O.qdot(c, t)[j]
        assert q[i, j] == pytest.approx(qo, rel=1e-12, abs=1e-13 * scale * phi), (m.reactions[j].equation, t)


def _random_states(m, ncells, seed):
    rng = np.random.default_rng(seed)
    T = rng.uniform(300.0, 3000.0, ncells)
    sc = 10.0 ** rng.uniform(-6.0, 0.0, (ncells, m.n_species)) * rng.uniform(1.0, 50.0, (ncells, 1))
    return sc, T


@pytest.mark.parametrize("mech,kernels", IMPLS)
def test_rates_and_jacobian_against_oracle(models, oracles, mech, kernels):
    m, O = models[mech], oracles[mech]
    ncells = 48 if m.n_species > 60 else 96
    sc, T = _random_states(m, ncells, 21)
    R = _handle(kernels)
    tsc, tT = _dev(sc), _dev(T)
    w = R.production_rate(tsc, tT).cpu().numpy()
    q = R.progress_rate(tsc, tT).cpu().numpy()
    n = m.n_species + 1
    for consP in (0, 1):
        J = R.ajacobian(tsc, tT, consP).cpu().numpy().reshape(ncells, n * n)
        for c in range(ncells):
            Jo = O.jacobian(sc[c], T[c], consP)
            assert np.abs(J[c] - Jo).max() <= 1e-11 * np.abs(Jo).max(), (c, consP)
    for c in range(ncells):
        wo, qo = O.wdot(sc[c], T[c]), O.qdot(sc[c], T[c])
        assert np.abs(w[c] - wo).max() <= 1e-11 * np.abs(wo).max(), c
        assert np.abs(q[c] - qo).max() <= 1e-11 * np.abs(qo).max(), c


def test_mechanism_entry_points_reject_bad_tensors():
    import torch
    R = _handle("LiDryer")
    sc = torch.zeros((4, 9), dtype=torch.float32, device="cuda:0")
    T = torch.full((4,), 1000.0, dtype=torch.float64, device="cuda:0")
    with pytest.raises(ValueError):
        R.production_rate(sc, T)
    with pytest.raises(ValueError):
        R.ajacobian(sc.double()[:3], T, 1)
