"""GPU parity on the larger mechanisms: dodecane_lu (53 species, 268 reactions)
and heptane_lu_88sk (88 species, 729 reactions, 718 of them irreversible).

Their per-cell matrices (Jacobian, LU factors, Krylov basis) do not fit a
lane's private memory and live in the per-lane HBM workspace; these tests run
every integrator through it.  The reference's GPU ReactorBDF stops at 54
species (ReactorBDF.cpp:52,147,331 abort for NUM_SPECIES >= 55): heptane is
compared with the CPU restatement of the same algorithm.

Workloads: dodecane cells drawn from the uneven-stiffness state (70% cold
fresh gas at 300-700 K, 25% near ignition at 900-1200 K and 1-40 atm, 5% burnt
at 2300 K, shuffled); heptane cells near ignition (800-1300 K, phi 0.5-2,
10-50 atm).  Criteria as in test_gpu_parity: |x - x_ref| <= 1e-6 |x_ref| +
1e-10 on T and Y_k, within the hull of the oracle's rounding ensemble for
the fixed-step BDF, exact per-cell step counts for the fixed-step and explicit
integrators.
"""
import numpy as np
import pytest

import oracle
from pelephysics_amd import workloads as W
from test_gpu_parity import _torch, _close, _compare_y, _run_1d, _oracle_1d, _hull_1d, _TY

pytestmark = pytest.mark.gpu

CASES = {"dodecane_lu": (W.dodecane_uneven, 64), "heptane_lu_88sk": (W.heptane_ignition, 24)}


def _case(models, mech, ncells=None, seed=W.SEED):
    gen, n = CASES[mech]
    m = models[mech]
    st = gen(m, ncells or n, seed=seed)
    return m, st, np.zeros((st.ncells, m.n_species)), np.zeros(st.ncells)


# banded cells measured on the CPU: dodecane 8 of 64, heptane 10 of 24 (near ignition)
@pytest.mark.parametrize("mech,max_banded", [("dodecane_lu", 0.15), ("heptane_lu_88sk", 0.45)])
def test_large_bdf_parity(models, oracles, mech, max_banded):
    """ReactorBDF defaults (BDF3 with TRPZ ramp, SGS-preconditioned GMRES with
    kspiters = N + 1) at dt = 1e-5 s."""
    torch = _torch()
    m, st, ysrc, esrc = _case(models, mech)
    ret, y, rX, R = _run_1d(torch, mech, "ReactorBDF", st, ysrc, esrc, 1e-5)
    oret, oy, orX, fc = _oracle_1d(oracles[mech], "bdf", st, ysrc, esrc, 1e-5)
    assert ret == oret == 1
    _compare_y(m.n_species + 1, y, oy, _hull_1d(models, mech, st, ysrc, esrc, 1e-5), max_banded)
    _close("rX", rX, orX)
    assert (R.cell_steps(st.ncells) == fc).all()


@pytest.mark.parametrize("mech,dt", [("dodecane_lu", 1e-8), ("heptane_lu_88sk", 1e-9)])
def test_large_rk64_parity(models, oracles, mech, dt):
    """ReactorRK64 defaults; the explicit scheme is stable on these mechanisms
    only for steps this short (at 1e-5 s both the oracle and the GPU diverge)."""
    torch = _torch()
    m, st, ysrc, esrc = _case(models, mech)
    ret, y, rX, R = _run_1d(torch, mech, "ReactorRK64", st, ysrc, esrc, dt)
    oret, oy, orX, fc = _oracle_1d(oracles[mech], "rk64", st, ysrc, esrc, dt)
    assert np.isfinite(oy).all()
    assert ret == oret
    assert (R.cell_steps(st.ncells) == fc).all()
    _compare_y(m.n_species + 1, y, oy)


@pytest.mark.parametrize("mech", ["dodecane_lu", "heptane_lu_88sk"])
@pytest.mark.parametrize("ident", ["ReactorBDFAdaptive", "ReactorBDFAdaptiveMixed", "ReactorBDFBatchedFP64"])
def test_large_adaptive_accuracy(models, oracles, mech, ident):
    """Adaptive forms (plain FP64; batched work queue with FP32 or FP64 LU) at
    rtol 1e-10 / atol 1e-16 against the oracle at rtol 1e-13 / atol 1e-20."""
    torch = _torch()
    m, st, ysrc, esrc = _case(models, mech)
    ret, y, rX, R = _run_1d(torch, mech, ident, st, ysrc, esrc, 1e-5, rtol=1e-10, atol=1e-16)
    assert R.stats()["nfailed"] == 0
    p = oracle.adapt_params(rtol=1e-13, atol=1e-20)
    _, oy, orX, _ = _oracle_1d(oracles[mech], "adaptive", st, ysrc, esrc, 1e-5, params=p)
    _compare_y(m.n_species + 1, y, oy)
    _close("rX", rX, orX)


def test_heptane_mixed_precision_chained(models, oracles):
    """FP32-LU Newton on heptane, 100 chained calls of 1e-5 s through ignition:
    per-component error against the FP64 oracle (rtol 1e-13) after the first
    call and after the last, each within the criterion.  The tolerance is
    rtol 1e-11 / atol 1e-20: over 100 calls through ignition the oracle's own
    FP64 result at rtol 1e-10 / atol 1e-16 drifts 3.7e-5 in T from the truth,
    at 1e-11 / 1e-20 by 1.3e-9."""
    torch = _torch()
    mech = "heptane_lu_88sk"
    m, st, _, _ = _case(models, mech, ncells=4, seed=3)
    n = m.n_species + 1
    from pelephysics_amd.reactor import ReactorBase
    R = ReactorBase.create("ReactorBDFAdaptiveMixed")
    R.init(2, st.ncells, mechanism=mech, rtol=1e-11, atol=1e-20)
    dev = torch.device("cuda:0")
    ty = torch.from_numpy(st.y_1d()).to(dev)
    tx = torch.from_numpy(st.rE.copy()).to(dev)
    ts = torch.zeros(st.ncells * m.n_species, dtype=torch.float64, device=dev)
    txs = torch.zeros(st.ncells, dtype=torch.float64, device=dev)
    O = oracles[mech]
    p = oracle.adapt_params(rtol=1e-13, atol=1e-20)
    oy, orX = st.y_1d(), st.rE.copy()
    zs, ze = np.zeros(st.ncells * m.n_species), np.zeros(st.ncells)
    errs = {}
    for call in range(100):
        R.react_1d(ty, ts, tx, txs, 1e-5)
        assert R.stats()["nfailed"] == 0
        O.react_1d("adaptive", oy, zs.copy(), orX, ze.copy(), 1e-5, params=p)
        if call % 10 == 9:
            print(f"call {call + 1}: GPU steps {R.stats()['nsteps']}", flush=True)
        if call in (0, 99):
            gT, gY = _TY(n, ty.cpu().numpy())
            rT, rY = _TY(n, oy)
            errs[call + 1] = (np.max(np.abs(gT - rT) / np.abs(rT)), np.max(np.abs(gY - rY) / (np.abs(rY) + 1e-4)))
            _compare_y(n, ty.cpu().numpy(), oy)
    print(f"heptane FP32-LU: max rel err T, Y after 1 call {errs[1]}, after 100 calls {errs[100]}")
    assert (oy.reshape(-1, n)[:, -1] > st.T + 1.0).any()  # some cells ignite over the 1 ms


@pytest.mark.parametrize("ident", ["ReactorBDF", "ReactorBDFAdaptive", "ReactorBDFAdaptiveMixed"])
def test_workspace_smaller_than_grid(models, ident):
    """A workspace of one 64-lane block: the lanes stride over (fixed and
    adaptive forms) or take from the queue (batched form) 200 cells, with
    results bitwise equal to those of the default workspace."""
    torch = _torch()
    mech = "dodecane_lu"
    m, st, ysrc, esrc = _case(models, mech, ncells=200)
    from pelephysics_amd.reactor import ReactorBase
    R = ReactorBase.create(ident)
    R.init(2, 1, mechanism=mech)
    rows = R.workspace_rows()
    assert rows > 0
    ode = dict(rtol=1e-10, atol=1e-16) if ident != "ReactorBDF" else {}
    _, y_full, rX_full, _ = _run_1d(torch, mech, ident, st, ysrc, esrc, 1e-5, **ode)
    _, y_one, rX_one, R1 = _run_1d(torch, mech, ident, st, ysrc, esrc, 1e-5,
                                   workspace_bytes=64 * rows * 8, **ode)
    assert R1.stats()["nfailed"] == 0
    assert np.array_equal(y_full, y_one) and np.array_equal(rX_full, rX_one)


@pytest.mark.parametrize("mech", ["dodecane_lu", "heptane_lu_88sk"])
def test_large_one_wave_build_tuned_service(models, oracles, mech):
    """The batched FP32-LU kernel built at 1 wave per SIMD with the tuned
    Jacobian/LU service (64 lanes or 8 trips) against the oracle."""
    torch = _torch()
    m, st, ysrc, esrc = _case(models, mech)
    ret, y, rX, R = _run_1d(torch, mech, "ReactorBDFAdaptiveMixed", st, ysrc, esrc, 1e-5, rtol=1e-10,
                            atol=1e-16, batched_waves_per_simd=1, batched_lu_batch=64, batched_lu_wait=8)
    assert R.stats()["nfailed"] == 0
    p = oracle.adapt_params(rtol=1e-13, atol=1e-20)
    _, oy, orX, _ = _oracle_1d(oracles[mech], "adaptive", st, ysrc, esrc, 1e-5, params=p)
    _compare_y(m.n_species + 1, y, oy)


def test_large_lu_service_parameters_bitwise(models):
    """With the matrices in the workspace too, the Jacobian/LU service
    parameters (every request served at once; 16 lanes or 4 trips; the default
    64 lanes or 32 trips) leave states and per-cell step counts bitwise
    unchanged."""
    torch = _torch()
    mech = "dodecane_lu"
    m, st, ysrc, esrc = _case(models, mech, ncells=128)
    runs = []
    for b, t in ((1, 1), (16, 4), (0, 0)):
        _, y, rX, R = _run_1d(torch, mech, "ReactorBDFAdaptiveMixed", st, ysrc, esrc, 1e-5, rtol=1e-10,
                              atol=1e-16, batched_lu_batch=b, batched_lu_wait=t)
        assert R.stats()["nfailed"] == 0
        runs.append((y, rX, R.cell_steps(st.ncells)))
    for y, rX, fc in runs[1:]:
        assert np.array_equal(y, runs[0][0]) and np.array_equal(rX, runs[0][1]) and np.array_equal(fc, runs[0][2])
