"""ReactorRK64 with the Manifold EOS's network model (manifold.model =
NeuralNet; the reference's Support/CMLM/fgm_net.pnn) on the GPU against the
oracle (tests/test_manifold_network.py pins the oracle's forward pass against
torch): 1-D and box forms on 1024 cells over the network's input range,
the same RK64 step counts and states within 1e-9 relative (the forward pass
sums in the same order; the device contracts multiply-adds)."""
import os

import numpy as np
import pytest

import oracle
from pelephysics_amd.manifold import ManifoldParm, NeuralNetwork
from test_gpu_parity import _torch
from test_manifold_network import INFO, P_CGS, PNN, points

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda:0")


def _setup(n, seed):
    net = NeuralNetwork.read(PNN, INFO)
    parm = ManifoldParm(net, P_CGS, True)
    rng = np.random.default_rng(seed)
    X = points(n, seed)
    rho = rng.uniform(8e-4, 1.5e-3, n)
    rY = np.column_stack([X * rho[:, None], rho])
    src = np.column_stack([rng.uniform(-1e-4, 1e-4, (n, 2)) * rho[:, None], np.zeros(n)])
    esrc = rng.uniform(-1e2, 1e2, n)
    return parm, oracle.Manifold(parm), rY, src, esrc


def _rk64(parm):
    from pelephysics_amd.reactor import ReactorBase
    R = ReactorBase.create("ReactorRK64")
    R.init(2, 1, mechanism="Manifold", manifold_dim=2)
    R.set_eos_parm(parm)
    return R


def _close(a, b, rel=1e-9):
    err = np.abs(a - b) / (np.abs(b) + 1e-300)
    assert np.all(np.abs(a - b) <= rel * np.abs(b) + 1e-30), float(err.max())


def test_network_1d_parity():
    _torch()
    n = 1024
    parm, O, rY, src, esrc = _setup(n, 21)
    y = np.column_stack([rY, np.full(n, 1200.0)])
    R = _rk64(parm)
    ty, ts, tx, txs = _dev(y.reshape(-1)), _dev(src.reshape(-1)), _dev(np.full(n, 3.0)), _dev(esrc)
    dt = 2e-5
    ret, _ = R.react_1d(ty, ts, tx, txs, dt)
    oy, ox = y.reshape(-1).copy(), np.full(n, 3.0)
    oret, ofc = O.rk64_1d(oy, np.ascontiguousarray(src.reshape(-1)), ox, esrc.copy(), dt)
    steps = R.cell_steps(n)
    print(f"network 1-D: avg steps {ret} (oracle {oret}), max {steps.max()}")
    assert ret == oret and np.array_equal(steps, ofc) and steps.max() > 1
    _close(ty.cpu().numpy(), oy)
    _close(tx.cpu().numpy(), ox)


def test_network_box_parity():
    torch = _torch()
    n = 1024
    parm, O, rY, src, esrc = _setup(n, 22)
    dt = 2e-5
    shape = (8, 8, 16)
    b = dict(rY=_dev(rY.T.reshape(3, *shape[::-1])), rYs=_dev(src.T.reshape(3, *shape[::-1])),
             T=_dev(np.zeros((1, *shape[::-1]))), rE=_dev(np.full((1, *shape[::-1]), 2.0)),
             rEs=_dev(esrc.reshape(1, *shape[::-1])), FC=_dev(np.zeros((1, *shape[::-1]))))
    R = _rk64(parm)
    ret, _ = R.react(b["rY"], b["rYs"], b["T"], b["rE"], b["rEs"], b["FC"], dt)
    orY, oT, orE = np.ascontiguousarray(rY.T).copy(), np.zeros(n), np.full(n, 2.0)
    oret, ofc = O.rk64_box(orY, np.ascontiguousarray(src.T), oT, orE, esrc.copy(), dt)
    assert ret == oret
    assert np.array_equal(b["FC"].cpu().numpy().reshape(-1), ofc)
    _close(b["rY"].cpu().numpy().reshape(3, n), orY)
    _close(b["T"].cpu().numpy().reshape(-1), oT)
    _close(b["rE"].cpu().numpy().reshape(-1), orE)
    # T comes from the network's T output at the final state
    Y = orY[:2] / orY[2]
    assert np.allclose(oT, [O.value(parm.idx_T, y) for y in Y.T], rtol=1e-14)


def test_network_refusals():
    _torch()
    from pelephysics_amd.reactor import ReactorBase, ReactorError
    net = NeuralNetwork.read(PNN, INFO)
    R = ReactorBase.create("ReactorRK64")
    R.init(2, 1, mechanism="Manifold")  # manifold_dim 1: the network has 2 inputs
    with pytest.raises(ReactorError, match="network inputs"):
        R.set_eos_parm(ManifoldParm(net, P_CGS, True))
