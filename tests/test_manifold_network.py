"""The Manifold EOS's network model (manifold.model = NeuralNet,
Source/Utility/BlackBoxFunction/NeuralNetHomerolled.H, NeuralNetModelDef.H,
NeuralNetLayerDef.H) on the host and on the oracle, with the reference's own
network Support/CMLM/fgm_net.pnn and its info file (tests/golden/, copied as
data; the reference's EosEval / TranEval inputs load them and hold no output
values):

* the reader: the 7 CMLM layers 2 -> 50 -> 50 -> 16, the info file's names,
  the EOS indices found from them, and the reader's refusals;
* the oracle's forward pass against torch.nn.functional's linear and
  leaky_relu and an affine BatchNorm (independent library routines), every
  output at random manifold points;
* the oracle's ReactorRK64 with the network: over a tiny step the state moves
  by dt x the network's SRC_xi outputs (checked against torch), the density
  by its external source and T by dt rhoE_src / rho.
"""
import os

import numpy as np
import pytest

import oracle
from pelephysics_amd.manifold import ManifoldParm, NeuralNetwork

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PNN = os.path.join(GOLDEN, "fgm_net.pnn")
INFO = os.path.join(GOLDEN, "fgm_net_info.txt")
P_CGS = 1013250.0


def torch_forward(net, X):
    import torch
    import torch.nn.functional as F
    x = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float64))
    for kind, nout, p in net.layers:
        p = torch.from_numpy(p)
        if kind == "Linear":
            nin = x.shape[1]
            x = F.linear(x, p[:nin * nout].reshape(nout, nin), p[nin * nout:])
        elif kind == "LeakyReLU":
            x = F.leaky_relu(x, float(p[0]))
        else:
            n = x.shape[1]
            x = x * p[:n] + p[n:]
    return x.numpy()


def points(n, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(-2.0, 2.0, n), rng.uniform(-0.1, 0.1, n)])


def test_reader_layers_names_indices():
    net = NeuralNetwork.read(PNN, INFO)
    assert [(k, n, p.size) for k, n, p in net.layers] == [
        ("Linear", 50, 150), ("LeakyReLU", 50, 1), ("BatchNorm1d", 50, 100), ("Linear", 50, 2550),
        ("LeakyReLU", 50, 1), ("BatchNorm1d", 50, 100), ("Linear", 16, 816)]
    assert net.model_name == "premixed" and net.dimnames == ["xi0", "xi1"] and net.nmanpar == 2
    assert net.varnames[-2:] == ["SRC_xi0", "SRC_xi1"] and len(net.varnames) == 16
    parm = ManifoldParm(net, P_CGS, True)
    assert (parm.idx_T, parm.idx_density, parm.idx_src) == (9, 10, [14, 15])
    assert ManifoldParm(net, P_CGS, True, density_lookup_type="log").idx_density == 11
    assert ManifoldParm(net, P_CGS, True, has_mani_src=False).idx_src == [-1, -1]


def test_reader_refusals(tmp_path):
    with pytest.raises(ValueError, match=".pnn extension"):
        NeuralNetwork.read(INFO, INFO)
    with pytest.raises(ValueError, match="CMLM"):
        NeuralNetwork.read(PNN, INFO, cmlm_net=False)
    raw = open(PNN, "rb").read()
    cut = tmp_path / "short.pnn"
    cut.write_bytes(raw[:len(raw) - (2 * 64 + 4 * (16 * 50 + 16))])  # the last Linear layer dropped
    with pytest.raises(ValueError, match="7 layers"):
        NeuralNetwork.read(str(cut), INFO)


def test_oracle_forward_against_torch():
    net = NeuralNetwork.read(PNN, INFO)
    O = oracle.Manifold(ManifoldParm(net, P_CGS, True))
    X = points(400, 1)
    ref = torch_forward(net, X)
    got = np.array([[O.value(v, x) for v in range(net.nvar)] for x in X])
    scale = np.abs(ref).max(0)
    assert np.all(np.abs(got - ref) <= 1e-12 * scale + 1e-13 * np.abs(ref))
    # the LeakyReLU kink matters: some hidden units are negative at these points
    h = torch_forward(NeuralNetwork(net.model_name, net.dimnames, net.varnames, 2, net.layers[:1]), X)
    assert (h < 0).any() and (h > 0).any()


def test_oracle_rk64_tiny_step_follows_the_network_sources():
    net = NeuralNetwork.read(PNN, INFO)
    parm = ManifoldParm(net, P_CGS, True)
    O = oracle.Manifold(parm)
    X = points(64, 2)
    rho = 1.2e-3
    src = torch_forward(net, X)[:, [14, 15]]
    n = X.shape[0]
    y = np.column_stack([X * rho, np.full(n, rho), np.full(n, 1500.0)])
    ysrc = np.zeros((n, 3))
    ysrc[:, 2] = 1e-7
    esrc = np.full(n, 5e3)
    dt = 1e-9
    y1, rX = y.reshape(-1).copy(), np.zeros(n)
    r, fc = O.rk64_1d(y1, ysrc.reshape(-1).copy(), rX, esrc.copy(), dt)
    y1 = y1.reshape(n, 4)
    rate = (y1[:, :2] - y[:, :2]) / dt
    assert np.all(np.abs(rate - src) <= 1e-5 * np.abs(src).max())
    assert np.allclose((y1[:, 2] - rho) / dt, 1e-7, rtol=1e-6)
    assert np.allclose((y1[:, 3] - 1500.0) / dt, 5e3 / rho, rtol=1e-6)
