"""The T-inversion term of the static RHS operation count
(pelephysics_amd/opcount.py TINV_ITERS) against the iterations the oracle's
GET_T_GIVEN_EY restatement counts inside the adaptive BDF on a sample of the
drm19 64^3 grid at the bench's step and tolerances (first and chained call),
and the counters themselves against a hand-checked case: a guess already
within the 1e-6 K tolerance takes one evaluation, a guess 300 K off takes
more."""
import numpy as np

import oracle
from pelephysics_amd import opcount
from pelephysics_amd import workloads as W


def test_tinv_counter_closed_cases(models):
    m = models["drm19"]
    O = oracle.Mech(m)
    st = W.drm19_grf(m, 4)
    Y = st.rY[0] / st.rY[0].sum()
    T0 = float(st.T[0])
    e, _ = O.e_cv_mass(T0, Y)
    with oracle.tinv_counts() as c:
        O.T_from_e(e, Y, T0)
    assert (c.calls, c.evals) == (1, 1)
    with oracle.tinv_counts() as c:
        T, _ = O.T_from_e(e, Y, T0 + 300.0)
    assert c.calls == 1 and c.evals >= 3 and abs(T - T0) < 1e-5
    with oracle.tinv_counts() as c:  # outside [90, 4000] K: extrapolated, not counted
        O.T_from_e(e * 50.0, Y, T0)
    assert c.calls == 0


def test_tinv_iters_matches_the_counted_mean(models):
    m = models["drm19"]
    O = oracle.Mech(m)
    sub = W.subsample(W.drm19_grf(m, 64), np.arange(0, 64 ** 3, 1024))
    p = oracle.adapt_params(rtol=1e-10, atol=1e-16)
    y, rX = sub.y_1d(), sub.rE.copy()
    z, src = np.zeros(sub.ncells), np.zeros(sub.ncells * m.n_species)
    means = []
    for _ in range(2):  # first call, then chained on the state it left
        with oracle.tinv_counts() as c:
            O.react_1d("adaptive", y, src, rX, z, 1e-5, params=p, nthreads=1)
        assert c.calls > 10 * sub.ncells
        means.append(c.mean)
    print("mean Newton evaluations per T inversion:", means)
    assert all(abs(x - opcount.TINV_ITERS) <= 0.5 for x in means)
