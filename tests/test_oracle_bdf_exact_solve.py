"""The oracle's fixed-step BDF with exact linear solves (gmres_precond = -1):
the map the discrete adjoint differentiates and the finite differences of the
adjoint tests go through.  Pinned against a numpy restatement of one call of
ReactorBDF::react (ReactorBDF.cpp:187-285: TRPZ ramp, one Newton iteration
per substep, history nm1/nm2 reset at every call) built from the oracle's RHS
and BDF Jacobian, with numpy.linalg.solve for the linear systems."""
import numpy as np
import pytest

import oracle
from pelephysics_amd import workloads as W

TC = {1: (1.0, -1.0, 0.0, 0.0), 2: (1.0, -1.3333, 0.3333, 0.0), 3: (1.0, -1.6363, 0.8181, -0.1818)}
FC = {1: (0.5, 0.5), 2: (0.6666, 0.0), 3: (0.5454, 0.0)}


def _numpy_call(O, u, rhoe, dt, nsub, scheme, rt=2):
    n = u.size
    h = dt / nsub
    first = 1 if scheme >= 2 else scheme
    change = scheme - 2
    soln_n, nm1, nm2 = u.copy(), u.copy(), u.copy()
    for st in range(nsub):
        ts = scheme if st > change else first
        t0, t1, t2, t3 = TC[ts]
        f0, f1 = FC[ts]
        tcur = st * h
        J = O.chem_jac(soln_n, rt)
        A = np.eye(n) * t0 / h - f0 * J
        f = O.fkernelspec(soln_n, rhoe, 0.0, np.zeros(n - 1), reactor_type=rt, dt_save=tcur)
        r = -(t0 * soln_n + t1 * soln_n + t2 * nm1 + t3 * nm2) / h + f0 * f + (f1 * f if ts == 1 else 0.0)
        new = soln_n + np.linalg.solve(A, r)
        nm2, nm1, soln_n = nm1, soln_n, new
    return soln_n


@pytest.mark.parametrize("nsub,scheme", [(1, 3), (4, 3), (3, 2)])
def test_exact_solve_map_matches_numpy(models, oracles, nsub, scheme):
    m = models["drm19"]
    st = W.subsample(W.drm19_grf(m, 4), np.array([5, 17, 40]))
    ns = m.n_species
    y = st.y_1d()
    rX = st.rE.copy()
    p = oracle.bdf_params(ns, nsubsteps=nsub, scheme=scheme, gmres_precond=-1)
    oracles["drm19"].react_1d("bdf", y, np.zeros(st.ncells * ns), rX, np.zeros(st.ncells), 1e-5, params=p)
    for c in range(st.ncells):
        ref = _numpy_call(oracles["drm19"], st.y_1d().reshape(-1, ns + 1)[c], st.rE[c], 1e-5, nsub, scheme)
        got = y.reshape(-1, ns + 1)[c]
        scale = np.abs(ref[:ns]).max()
        assert np.abs(got[:ns] - ref[:ns]).max() <= 1e-9 * scale
        assert abs(got[-1] - ref[-1]) <= 1e-9 * ref[-1]


def test_exact_solve_conserves_mass(models, oracles):
    """sum_k rhoY_k is an invariant of the map (1^T A = 1^T/dt on the species rows)."""
    m = models["drm19"]
    st = W.subsample(W.drm19_fresh(m, 16), np.arange(16))
    ns = m.n_species
    y = st.y_1d()
    rX = st.rE.copy()
    for _ in range(10):
        oracles["drm19"].react_1d("bdf", y, np.zeros(st.ncells * ns), rX, np.zeros(st.ncells), 1e-5,
                                  params=oracle.bdf_params(ns, gmres_precond=-1))
    s0 = st.rY.sum(1)
    s1 = y.reshape(-1, ns + 1)[:, :ns].sum(1)
    assert np.abs(s1 - s0).max() <= 1e-13 * s0.max()
