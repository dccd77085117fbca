"""A PeleLMeX-style plotfile (MKS) read into the reactor state and integrated
by the box form on the GPU, against the oracle on the same CGS arrays
(the ReactEval path: Testing/Exec/ReactEval/main.cpp:310-387, then react())."""
import numpy as np
import pytest

import oracle
from pelephysics_amd import pltfile
from pelephysics_amd import workloads as W
from test_gpu_parity import _torch, _close

pytestmark = pytest.mark.gpu


def test_react_from_plotfile(models, oracles, tmp_path):
    torch = _torch()
    from pelephysics_amd.reactor import ReactorBase
    m = models["drm19"]
    ns = m.n_species
    n = 8
    st = W.drm19_grf(m, n)  # 512 cells, i fastest
    rng = np.random.default_rng(7)
    rYs = 1e-3 * rng.standard_normal(st.rY.shape) * st.rY
    rEs = 1e-3 * rng.standard_normal(st.ncells) * np.abs(st.rE)

    def grid(a):  # (ncells, k) -> (k, nz, ny, nx)
        return np.ascontiguousarray(a.T).reshape(-1, n, n, n)

    sp = [s.name for s in m.species]
    names = ["density"] + [f"rhoY({s})" for s in sp] + ["Temp", "rhoH"] + [f"F_rhoY({s})" for s in sp] + ["F_rhoH"]
    data = np.concatenate([grid(st.rY.sum(1, keepdims=True)) * 1e3, grid(st.rY) * 1e3, grid(st.T[:, None]),
                           grid(st.rE[:, None]) / 10.0, grid(rYs) * 1e3, grid(rEs[:, None]) / 10.0])
    path = str(tmp_path / "plt_drm19")
    pltfile.write(path, names, data, max_grid_size=4)
    s = pltfile.reactor_state(pltfile.PlotFile(path), sp)
    assert np.allclose(s["rY"], grid(st.rY), rtol=1e-15, atol=0)

    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(v.copy()).to(dev) for k, v in s.items()}
    FC = torch.zeros_like(t["T"])
    R = ReactorBase.create("ReactorBDFAdaptiveMixed")
    R.init(2, st.ncells, mechanism="drm19", rtol=1e-10, atol=1e-16)
    R.react(t["rY"], t["rYsrc"], t["T"], t["rE"], t["rEsrc"], FC, 1e-5)
    assert R.stats()["nfailed"] == 0

    oY, oT, orE = s["rY"].reshape(ns, -1).copy(), s["T"].reshape(-1).copy(), s["rE"].reshape(-1).copy()
    oracles["drm19"].react_box("adaptive", oY, s["rYsrc"].reshape(ns, -1).copy(), oT, orE,
                               s["rEsrc"].reshape(-1).copy(), 1e-5, params=oracle.adapt_params(rtol=1e-13, atol=1e-20))
    gY = t["rY"].cpu().numpy().reshape(ns, -1)
    _close("T", t["T"].cpu().numpy().reshape(-1), oT)
    _close("Y", (gY / gY.sum(0)).T, (oY / oY.sum(0)).T)
    _close("rE", t["rE"].cpu().numpy().reshape(-1), orE)
