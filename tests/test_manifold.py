"""The Manifold EOS path of ReactorRK64 on the CPU: the table file, the EOS
parameters, and the oracle's lookups and integration.

Reference: Source/Eos/Manifold.H (RY2RRinvY takes rho from the last species;
RHY2T / REY2T / RTY2WDOT / PYT2R are table lookups), Source/Utility/
BlackBoxFunction/Table.H (read_table; multilinear interpolation with indices
clamped to the grid and weights to [0, 1]), Source/Eos/EosParams.H:84-160
(indices from the manifold.* keys), ReactorUtils.H:358-415 (fKernelSpec with
USE_MANIFOLD_EOS), ReactorRK64.cpp (the integrator).

Pins:

* the table fixture (tests/golden/autoignition.ctb, the reference's
  Support/CMLM/autoignition.ctb, the table of its Manifold ignition-delay
  test) reads to the layout read_table expects and writes back byte for byte;
* lookups reproduce a function that is linear in each coordinate exactly
  (the closed form of multilinear interpolation), return the tabulated value
  at every node, and hold the boundary value outside the grid;
* a constant source integrates exactly in the closed number of RK64 steps, a
  linear source to the exponential closed form;
* the reference's Manifold ignition-delay known-answer test
  (.github/workflows/ci.yml:279-281: check_ignDelay.py 0.0766 10, with
  inputs/inputs.0d_manifold: RK64, h-reactor, T0 = 600 K, p = 5 atm, calls
  of 1 s / 2000, then 1.1 x the estimate / 4000, delay when T exceeds T0 +
  400 K) on the oracle.
"""
import filecmp
import os

import numpy as np
import pytest

import oracle
from pelephysics_amd import workloads as W
from pelephysics_amd.manifold import ManifoldParm, TabulatedFunction

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CTB = os.path.join(GOLDEN, "autoignition.ctb")
T_IGN, FACTOR = 0.0766, 10.0  # ci.yml:281


def test_table_fixture_layout_and_round_trip(tmp_path):
    t = TabulatedFunction.read(CTB)
    assert t.ndim == 1 and t.dimnames == ["PROG"] and t.dim_lengths == [1119]
    assert t.model_name == "AUTOIGNITION"
    assert t.varnames[:5] == ["RHO", "T", "DIFF", "VISC", "SRC_PROG"] and t.nvar == 15
    assert np.all(np.diff(t.grids[0]) > 0) and t.grids[0][0] == 0.0
    T = t.values[t.var_index("T")]
    assert T[0] == 600.0 and 1126.0 < T.max() < 1127.0
    out = tmp_path / "copy.ctb"
    t.write(str(out))
    assert filecmp.cmp(str(out), CTB, shallow=False)


def _synthetic(tmp_path, dims, names, fn, varnames=None, extra=None):
    """A table of f(x) = fn(x) over the given grids, written and read back."""
    mesh = np.meshgrid(*dims[::-1], indexing="ij")[::-1]  # first dimension fastest
    vals = {"RHO": np.full(mesh[0].shape, 1.5e-3), "T": 300.0 + 1000.0 * mesh[0], "F": fn(*mesh)}
    vals.update(extra or {})
    varnames = varnames or list(vals)
    t = TabulatedFunction(names, [np.asarray(g, float) for g in dims], varnames,
                          np.stack([vals[v].reshape(-1) for v in varnames]), "TEST")
    p = tmp_path / "t.ctb"
    t.write(str(p))
    return TabulatedFunction.read(str(p))


def test_parameter_indices(tmp_path):
    """EosParams.H:137-154: the source of "Y-<name>" is "SRC_<name>", of any
    other parameter p "SRC_p"; missing sources are -1 (zero); without
    has_mani_src every source is -1."""
    g = [np.linspace(0, 1, 3), np.linspace(0, 1, 4)]
    t = _synthetic(tmp_path, g, ["Y-CO2", "ZMIX"], lambda a, b: a + b,
                   varnames=["T", "RHO", "lnRHO", "F", "SRC_CO2"],
                   extra={"lnRHO": np.zeros((4, 3)), "SRC_CO2": np.ones((4, 3))})
    p = ManifoldParm(t, 1e6, True)
    assert (p.idx_density, p.idx_T, p.idx_src) == (1, 0, [4, -1])
    assert ManifoldParm(t, 1e6, True, density_lookup_type="log").idx_density == 2
    assert ManifoldParm(t, 1e6, True, has_mani_src=False).idx_src == [-1, -1]
    with pytest.raises(KeyError):
        ManifoldParm(t, 1e6, True, density_lookup_type="inverse")
    with pytest.raises(ValueError):
        ManifoldParm(t, 1e6, True, density_lookup_type="cubic")


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_lookup_closed_forms(tmp_path, ndim):
    rng = np.random.default_rng(7 + ndim)
    dims = [np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0, 1, 5 + 2 * k)])) for k in range(ndim)]
    coef = rng.normal(size=2 ** ndim)

    def f(*x):  # sum over subsets S of coef_S prod_{k in S} x_k: linear in each coordinate
        out = 0.0
        for s in range(2 ** ndim):
            term = coef[s]
            for k in range(ndim):
                if s >> k & 1:
                    term = term * x[k]
            out = out + term
        return out

    names = ["PROG", "ZMIX", "ZVAR"][:ndim]
    t = _synthetic(tmp_path, dims, names, f)
    O = oracle.Manifold(ManifoldParm(t, 1e6, True))
    iF = t.var_index("F")
    for _ in range(200):
        x = rng.uniform(0, 1, ndim)
        assert O.value(iF, x) == pytest.approx(f(*x), rel=1e-13, abs=1e-13)
    # every node: the tabulated value (to the rounding of the weight
    # (x_{i+1} - x) * (1 / (x_{i+1} - x_i)), which can miss 1 by an ulp)
    for idx in np.ndindex(*[len(g) for g in dims]):
        x = np.array([dims[k][idx[k]] for k in range(ndim)])
        flat = sum(idx[k] * int(np.prod([len(g) for g in dims[:k]])) for k in range(ndim))
        assert O.value(iF, x) == pytest.approx(t.values[iF][flat], rel=1e-15, abs=1e-15)
    # outside the grid: the weights are clamped, the boundary value is held
    x = rng.uniform(0, 1, ndim)
    lo, hi = x.copy(), x.copy()
    lo[0], hi[0] = -0.5, 2.0
    e0, e1 = x.copy(), x.copy()
    e0[0], e1[0] = 0.0, 1.0
    assert O.value(iF, lo) == pytest.approx(O.value(iF, e0), rel=1e-15, abs=1e-15)
    assert O.value(iF, hi) == pytest.approx(O.value(iF, e1), rel=1e-15, abs=1e-15)


def test_density_lookup_types(tmp_path):
    g = [np.linspace(0, 1, 11)]
    rho = lambda x: 1e-3 * (1.0 + x)
    extra = {"lnRHO": np.log(rho(g[0])), "invRHO": 1.0 / rho(g[0])}
    t = _synthetic(tmp_path, g, ["PROG"], lambda x: x, varnames=["RHO", "T", "F", "lnRHO", "invRHO"],
                   extra=dict(extra, RHO=rho(g[0])))
    for kind in ("linear", "log", "inverse"):
        O = oracle.Manifold(ManifoldParm(t, 1e6, True, density_lookup_type=kind))
        for x in (0.0, 0.3, 1.0):  # on nodes every transform returns rho itself
            assert O.density(np.array([x, 1.0])) == pytest.approx(rho(x), rel=1e-14)


def _one_d_table(tmp_path, src):
    g = [np.linspace(0.0, 1.0, 101)]
    return _synthetic(tmp_path, g, ["PROG"], lambda x: x, varnames=["RHO", "T", "SRC_PROG"],
                      extra={"SRC_PROG": src(g[0])})


def test_rk64_constant_source_closed_form(tmp_path):
    """d(rho PV)/dt = c: exact after 6 steps (dt/10 first; a zero error
    estimate grows the step to its cap dt/5; the last step is clipped)."""
    c, rho, dt = 3.0, 2.0, 0.01
    O = oracle.Manifold(ManifoldParm(_one_d_table(tmp_path, lambda x: c + 0 * x), 1e6, True))
    y = np.array([0.1 * rho, rho, 900.0])
    rX, rXs = np.array([0.0]), np.array([5.0])
    r, fc = O.rk64_1d(y, np.zeros(2), rX, rXs, dt)
    assert fc[0] == 6 and r == 6
    assert y[0] == pytest.approx(0.1 * rho + c * dt, rel=1e-14)
    assert y[1] == rho
    assert y[2] == pytest.approx(900.0 + 5.0 / rho * dt, rel=1e-14)  # dT/dt = rhoE_src / rho
    assert rX[0] == pytest.approx(5.0 * dt, rel=1e-15)


def test_rk64_linear_source_closed_form(tmp_path):
    """SRC = a + b PV with rho constant: PV(t) = (PV0 + a/b) e^{b t / rho} - a/b,
    which the grid interpolates exactly; the box form returns T from the table."""
    a, b, rho, dt = 0.5, 4.0, 1.25, 0.05
    parm = ManifoldParm(_one_d_table(tmp_path, lambda x: a + b * x), 1e6, True)
    O = oracle.Manifold(parm)
    pv0 = 0.05
    rY = np.array([[pv0 * rho], [rho]])
    T, rE = np.array([0.0]), np.array([0.0])
    r, fc = O.rk64_box(rY, np.zeros((2, 1)), T, rE, np.zeros(1), dt, params=oracle.rk64_params(atol=1e-13))
    pv = (pv0 + a / b) * np.exp(b * dt / rho) - a / b
    assert rY[0, 0] / rho == pytest.approx(pv, rel=1e-9)
    assert T[0] == pytest.approx(300.0 + 1000.0 * rY[0, 0] / rho, rel=1e-14)
    assert fc[0] == r > 6


def test_manifold_ignition_delay_kat():
    """The reference's Manifold ignition-delay test on the oracle: the delay
    as its log prints it (3 significant digits) within 10 call lengths of
    0.0766 s."""
    parm = ManifoldParm.from_file(CTB, 5066250.0, True)
    O = oracle.Manifold(parm)
    rho = O.density(np.array([0.0, 1.0]))  # PYT2R at Y = (PV 0, rho 1) (GPU_misc.H:67-68)
    st = {}

    def advance(start, dtc, k):
        if start:
            st["rY"], st["T"], st["rE"] = np.array([[0.0], [rho]]), np.array([600.0]), np.array([0.0])
        Ts = []
        for _ in range(k):
            O.rk64_box(st["rY"], np.zeros((2, 1)), st["T"], st["rE"], np.zeros(1), dtc)
            Ts.append(st["T"][0])
        return Ts

    t_ign, unc = W.ignition_delay(advance, 600.0, dt=1.0, ndt=2000)
    t_log, unc_log = float(f"{t_ign:.3g}"), float(f"{unc:.3g}")
    print(f"ignition delay {t_ign:.6g} s +/- {unc:.3g} s (reference 0.0766 s)")
    assert abs(t_log - T_IGN) <= FACTOR * unc_log


def test_non_rk64_reactors_refuse_the_manifold():
    from pelephysics_amd.reactor import ReactorBase, ReactorError
    parm = ManifoldParm.from_file(CTB, 5066250.0, True)
    for ident in ("ReactorBDF", "ReactorBDFAdaptive", "ReactorCvode", "ReactorBDFAdaptiveMixed"):
        with pytest.raises(ReactorError, match="Manifold EOS only supported with ReactorRK64"):
            ReactorBase.create(ident).set_eos_parm(parm)
