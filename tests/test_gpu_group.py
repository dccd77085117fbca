"""The batched adaptive BDF with a group of lanes per cell
(batched_cells_per_wave = 8, 4, 2, 1: 8, 16, 32 or 64 lanes per cell;
group_bdf.h) against the tight FP64 oracle, at the benchmark tolerances
(GPU rtol 1e-10 / atol 1e-16, LiDryer 1e-12 / 1e-18; oracle rtol 1e-13 /
atol 1e-20):

* every cell meets |x - x_ref| <= 1e-6 |x_ref| + 1e-10 on T and Y_k, with
  the FP32 and the FP64 Newton-matrix LU, on every mechanism and group size
  the kernels are built for;
* two calls on the same input give the same state bit for bit, whatever
  order the queue visits the cells in (the second call is cost-ordered);
* the per-cell work is that of the one-cell-per-lane kernel up to rounding
  (same step / order / Newton logic);
* box calls run through flat state and meet the criterion against the
  oracle's box form; group sizes a mechanism has no kernel for are rejected
  at init.
"""
import numpy as np
import pytest

import oracle
from pelephysics_amd import workloads as W
from test_gpu_parity import _torch, _compare_y, _oracle_1d

pytestmark = pytest.mark.gpu

CASES = [("drm19_table", 8), ("drm19_table", 4), ("drm19_table", 2), ("LiDryer_table", 8), ("LiDryer_table", 4),
         ("dodecane_lu", 2), ("dodecane_lu", 1), ("heptane_lu_88sk", 2), ("heptane_lu_88sk", 1)]


def _cells(models, mech):
    m = models[mech]
    if mech == "drm19":
        return m, W.subsample(W.drm19_grf(m, 64), np.arange(0, 64 ** 3, 997))
    if mech == "LiDryer":
        return m, W.lidryer_h2air(m, 128)
    if mech == "dodecane_lu":
        return m, W.dodecane_uneven(m, 64)
    return m, W.heptane_ignition(m, 24)


# LiDryer's 1000-1400 K cells ignite within the call and amplify the local
# errors: at rtol 1e-10 every form, the plain FP64 one included, misses the
# criterion by the same 18x, so LiDryer runs at rtol 1e-12 / atol 1e-18
TOLS = {"LiDryer": (1e-12, 1e-18)}


def _run(torch, kernels, ident, st, m, calls=2, **ode):
    from pelephysics_amd.reactor import ReactorBase
    R = ReactorBase.create(ident)
    rtol, atol = TOLS.get(kernels.replace("_table", ""), (1e-10, 1e-16))
    R.init(2, st.ncells, mechanism=kernels, rtol=rtol, atol=atol, **ode)
    dev = torch.device("cuda:0")
    ty = torch.from_numpy(st.y_1d()).to(dev)
    ts = torch.zeros(st.ncells * m.n_species, dtype=torch.float64, device=dev)
    tx = torch.from_numpy(st.rE.copy()).to(dev)
    txs = torch.zeros(st.ncells, dtype=torch.float64, device=dev)
    outs = []
    for _ in range(calls):
        ty.copy_(torch.from_numpy(st.y_1d()))
        tx.copy_(torch.from_numpy(st.rE.copy()))
        R.react_1d(ty, ts, tx, txs, 1e-5)
        torch.cuda.synchronize()
        s = R.stats()
        assert s["nfailed"] == 0
        outs.append((ty.cpu().numpy(), tx.cpu().numpy(), R.cell_steps(st.ncells), s))
    return outs


_ORACLE = {}


def _reference(oracles, mech, st, m):
    if mech not in _ORACLE:
        z = np.zeros((st.ncells, m.n_species))
        _ORACLE[mech] = _oracle_1d(oracles[mech], "adaptive", st, z, np.zeros(st.ncells), 1e-5,
                                   params=oracle.adapt_params(rtol=1e-13, atol=1e-20))[1]
    return _ORACLE[mech]


@pytest.mark.parametrize("ident", ["ReactorBDFAdaptiveMixed", "ReactorBDFBatchedFP64"])
@pytest.mark.parametrize("kernels,cpw", CASES)
def test_group_kernel_accuracy(models, oracles, kernels, cpw, ident):
    torch = _torch()
    mech = kernels.replace("_table", "")
    print(f"group kernel {kernels} cells_per_wave={cpw} {ident}", flush=True)
    m, st = _cells(models, mech)
    outs = _run(torch, kernels, ident, st, m, batched_cells_per_wave=cpw)
    oy = _reference(oracles, mech, st, m)
    _compare_y(m.n_species + 1, outs[0][0], oy)
    # the cost-ordered second call changes no cell's result
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][2], outs[1][2])
    # the same algorithm as the one-cell-per-lane kernel: total work within 5%
    # (the LU rounding differs, which can move a Newton or error-test decision)
    lane = _run(torch, kernels, ident, st, m, calls=1)[0][3]
    s = outs[0][3]
    for k in ("nsteps", "nrhs", "nlu"):
        assert abs(s[k] - lane[k]) <= 0.05 * lane[k], (k, s[k], lane[k])


def test_group_box_route_and_validation(models, oracles):
    """A box call (through flat state) against the oracle's box form, and the
    configurations that have no group kernel."""
    torch = _torch()
    from pelephysics_amd.reactor import ReactorBase, ReactorError
    m = models["drm19"]
    st = W.drm19_grf(m, 8)
    n = 8
    dev = torch.device("cuda:0")

    def soa(a, nc):
        return torch.from_numpy(np.ascontiguousarray(a.T.reshape(nc, n, n, n))).to(dev)

    R = ReactorBase.create("ReactorBDFAdaptiveMixed")
    R.init(2, st.ncells, mechanism="drm19_table", rtol=1e-10, atol=1e-16, batched_cells_per_wave=4)
    rY, T, rE = soa(st.rY, m.n_species), soa(st.T[:, None], 1), soa(st.rE[:, None], 1)
    rYs, rEs, FC = torch.zeros_like(rY), torch.zeros_like(rE), torch.zeros_like(T)
    R.react(rY, rYs, T, rE, rEs, FC, 1e-5)
    assert R.stats()["nfailed"] == 0
    oY = np.ascontiguousarray(st.rY.T).copy()
    oT, orE = st.T.copy(), st.rE.copy()
    oracles["drm19"].react_box("adaptive", oY, np.zeros_like(oY), oT, orE, np.zeros(st.ncells), 1e-5,
                               params=oracle.adapt_params(rtol=1e-13, atol=1e-20))
    gY = rY.reshape(m.n_species, -1).cpu().numpy().T
    gT = T.reshape(-1).cpu().numpy()
    tol = lambda ref: 1e-6 * np.abs(ref) + 1e-10
    assert (np.abs(gT - oT) <= tol(oT)).all()
    Yg, Yr = gY / gY.sum(1, keepdims=True), oY.T / oY.T.sum(1, keepdims=True)
    assert (np.abs(Yg - Yr) <= tol(Yr)).all()
    assert (FC.reshape(-1).cpu().numpy() == R.cell_steps(st.ncells)).all()
    for bad in (dict(batched_cells_per_wave=3), dict(batched_cells_per_wave=4, batched_box_direct=1)):
        with pytest.raises(ReactorError):
            ReactorBase.create("ReactorBDFAdaptiveMixed").init(2, 8, mechanism="drm19_table", **bad)
    # the generated straight-line kernels have no group form
    with pytest.raises(ReactorError):
        ReactorBase.create("ReactorBDFAdaptiveMixed").init(2, 8, mechanism="drm19", batched_cells_per_wave=4)
    # nor does a one-lane-per-cell integrator
    with pytest.raises(ReactorError):
        ReactorBase.create("ReactorBDFAdaptive").init(2, 8, mechanism="drm19_table", batched_cells_per_wave=4)
