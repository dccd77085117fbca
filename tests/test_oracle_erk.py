"""ReactorArkode's explicit Runge-Kutta integration on the oracle
(ppc_oracle.c erk_core): the tables of pelephysics_amd/erk_tables.py stepped
with an embedded error estimate and ARKODE's published adaptivity.

SUNDIALS is absent from the reference tree (Submodules/sundials is empty), so
the stepping is pinned by what the mathematics fixes rather than against
ARKODE itself:

* fixed steps on y' = -y converge at each table's order q;
* with error control, the global error on a linear system follows rtol;
* the RHS count closed form: 2 (f(y0) and the initial-step probe) plus s - 1
  per attempt plus, for tables without first-same-as-last, one more per
  accepted step but the last;
* the controllers' fixed points: on y' = -y with relative tolerance the step
  size settles where the factor safety x (1.5 e)^-(k1 - k2 + k3)/k lies in
  the no-change band [1, 1.5] (PID: k1 - k2 + k3 = 0.47, PI: 0.49, I: 1,
  explicit Gustafsson: k1 = 0.367), so the last error norm lies in a band
  fixed by the controller's exponents (a wrong sign or exponent lands
  elsewhere), and each controller gives its own step count;
* chemistry: the drm19 and LiDryer cells against the adaptive BDF at rtol
  1e-12 (parity unpinned against SUNDIALS' ARKODE itself).
"""
import numpy as np
import pytest
from scipy.linalg import expm

import oracle
from pelephysics_amd import erk_tables as E
from pelephysics_amd import workloads as W

A2 = np.array([[-2.0, 1.0], [0.5, -3.0]])
Y0 = np.array([1.0, 0.5])


@pytest.mark.parametrize("code", sorted(E.TABLES))
def test_fixed_step_order(code):
    """y' = -y over [0, 1]: halving the step divides the error by 2^q."""
    q = E.dense(code)[2]
    hs = (0.125, 0.0625) if q <= 4 else ((0.25, 0.125) if q <= 6 else (0.5, 0.25))  # exact sums of steps
    err = []
    for h in hs:
        r, y, st = oracle.erk_linear(np.array([[-1.0]]), np.array([1.0]), 1.0, code, fixed_h=h)
        assert r > 0 and st["nst"] == round(1.0 / h)
        err.append(abs(y[0] - np.exp(-1.0)))
    assert np.log2(err[0] / err[1]) == pytest.approx(q, abs=0.4), err


@pytest.mark.parametrize("code", sorted(E.TABLES))
def test_adaptive_error_follows_rtol(code):
    ex = expm(A2 * 2.0) @ Y0
    errs = []
    for rtol in (1e-6, 1e-9):
        r, y, st = oracle.erk_linear(A2, Y0, 2.0, code, rtol=rtol, atol=1e-14, max_steps=10 ** 6)
        assert st["status"] == 0
        errs.append(np.abs(y - ex).max() / np.abs(ex).max())
        assert errs[-1] <= 50 * rtol
    assert errs[1] < errs[0]


@pytest.mark.parametrize("code", sorted(E.TABLES))
def test_rhs_count_closed_form(code):
    s = E.dense(code)[1]
    r, y, st = oracle.erk_linear(A2 * 40.0, Y0, 1.0, code, rtol=1e-7, atol=1e-12, max_steps=10 ** 6)
    attempts = st["nst"] + st["nrej"]
    extra = 0 if E.fsal(code) else st["nst"] - 1
    assert st["nfe"] == r == 2 + (s - 1) * attempts + extra


@pytest.mark.parametrize("controller,expo", [(0, 0.58 - 0.21 + 0.1), (1, 0.8 - 0.31), (2, 1.0), (3, 0.367)])
def test_controller_fixed_point_band(controller, expo):
    code = 52  # Dormand-Prince: embedding order 4, k = 5
    k = E.dense(code)[3] + 1.0
    r, y, st = oracle.erk_linear(np.array([[-1.0]]), np.array([1.0]), 30.0, code, rtol=1e-7, atol=1e-300,
                                 controller=controller, max_steps=10 ** 6)
    assert st["status"] == 0 and y[0] == pytest.approx(np.exp(-30.0), rel=1e-5)
    lo = (0.96 / 1.5) ** (k / expo) / 1.5
    hi = 0.96 ** (k / expo) / 1.5
    print(f"controller {controller}: steps {st['nst']}, rejections {st['nrej']}, last norm {st['last_dsm']:.4g} "
          f"in [{lo:.4g}, {hi:.4g}]")
    assert lo * (1 - 1e-9) <= st["last_dsm"] <= hi * (1 + 1e-9)


def test_controllers_differ():
    steps = set()
    for ctrl in range(4):
        r, y, st = oracle.erk_linear(A2 * 40.0, Y0, 1.0, 40, rtol=1e-8, atol=1e-12, controller=ctrl,
                                     max_steps=10 ** 6)
        steps.add((st["nst"], st["nrej"]))
    assert len(steps) == 4


@pytest.mark.parametrize("mech,code,dt", [("LiDryer", 40, 1e-5), ("LiDryer", 52, 1e-5), ("drm19", 40, 2e-6),
                                          ("drm19", 60, 2e-6), ("drm19", 80, 2e-6)])
def test_chemistry_against_adaptive_bdf(models, oracles, mech, code, dt):
    m, O = models[mech], oracles[mech]
    n = 48
    st = W.lidryer_h2air(m, n, seed=3) if mech == "LiDryer" else W.drm19_fresh(m, n, seed=3)
    ns = m.n_species
    y, x = st.y_1d().copy(), st.rE.copy()
    r, stats = O.react_1d("erk", y, np.zeros(n * ns), x, np.zeros(n), dt,
                          params=dict(code=code, rtol=1e-10, atol=1e-20, max_steps=10 ** 6))
    assert r > 0 and all(s["status"] == 0 for s in stats)
    yb, xb = st.y_1d().copy(), st.rE.copy()
    O.react_1d("adaptive", yb, np.zeros(n * ns), xb, np.zeros(n), dt, params=oracle.adapt_params(rtol=1e-13,
                                                                                                  atol=1e-20))
    a, b = y.reshape(n, ns + 1), yb.reshape(n, ns + 1)
    T, Tb = a[:, -1], b[:, -1]
    Y, Yb = a[:, :-1] / a[:, :-1].sum(1, keepdims=True), b[:, :-1] / b[:, :-1].sum(1, keepdims=True)
    assert np.all(np.abs(T - Tb) <= 1e-6 * Tb)
    assert np.all(np.abs(Y - Yb) <= 1e-6 * np.abs(Yb) + 1e-10)
    assert r == int(sum(s["nfe"] for s in stats) / n)
