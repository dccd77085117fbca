"""Chained 1-D calls replayed from a HIP graph (pp_chem_react_1d_chained)
against the same number of ordinary calls: every integrated value and step
count must be bitwise equal (the per-cell integration does not depend on the
queue order).  The graph is replayed again after eager work on the same
arrays, as a flow solver would interleave other kernels between chemistry
steps."""
import numpy as np
import pytest

from pelephysics_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ident", ["ReactorBDFAdaptiveMixed", "ReactorBDF", "ReactorRK64"])
def test_chained_graph_equals_eager(models, ident):
    import torch
    assert torch.cuda.is_available()
    from pelephysics_amd.reactor import ReactorBase
    m = models["LiDryer"]
    st = W.lidryer_h2air(m, 300, seed=21)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(21)
    src = 10.0 * rng.standard_normal((st.ncells, m.n_species)) * st.rY  # relative rates of 10/s
    iN2 = m.species_index("N2")
    src[:, iN2] -= src.sum(1)                                           # no net mass source
    ysrc = torch.from_numpy(src.reshape(-1).copy()).to(dev)
    esrc = torch.from_numpy(10.0 * rng.standard_normal(st.ncells) * np.abs(st.rE)).to(dev)
    ode = dict(rtol=1e-10, atol=1e-16) if "Adaptive" in ident else {}
    # the one-Newton-iteration fixed-step BDF stays bounded on these igniting
    # cells over 12 calls of 5e-7 s (not of 2e-6 s)
    dt = 5e-7 if ident == "ReactorBDF" else 2e-6

    def fresh():
        return torch.from_numpy(st.y_1d()).to(dev), torch.from_numpy(st.rE.copy()).to(dev)

    RA, RB = ReactorBase.create(ident), ReactorBase.create(ident)
    RA.init(2, st.ncells, mechanism="LiDryer", **ode)
    RB.init(2, st.ncells, mechanism="LiDryer", **ode)
    yA, xA = fresh()
    yB, xB = fresh()
    for _ in range(8):
        RA.react_1d(yA, ysrc, xA, esrc, dt)
    RB.react_1d_chained(yB, ysrc, xB, esrc, dt, 8)
    assert torch.isfinite(yA).all()
    assert torch.equal(yA, yB) and torch.equal(xA, xB)
    assert np.array_equal(RA.cell_steps(st.ncells), RB.cell_steps(st.ncells))
    # eager work on both, then a replay of the cached graph
    RA.react_1d(yA, ysrc, xA, esrc, dt)
    RB.react_1d(yB, ysrc, xB, esrc, dt)
    yB.add_(0.0)
    for _ in range(3):
        RA.react_1d(yA, ysrc, xA, esrc, dt)
    RB.react_1d_chained(yB, ysrc, xB, esrc, dt, 3)
    assert torch.equal(yA, yB) and torch.equal(xA, xB)
    assert np.array_equal(RA.cell_steps(st.ncells), RB.cell_steps(st.ncells))
    assert torch.isfinite(yB).all()
