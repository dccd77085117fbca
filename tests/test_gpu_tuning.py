"""A reactor initialised from the committed tuning table runs the tuned
kernels and launch shape, with results inside the criterion of the default
ones (the implementations differ in rounding only)."""
import numpy as np
import pytest

from pelephysics_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_tuned_init_1d(models):
    import torch
    assert torch.cuda.is_available()
    from pelephysics_amd.reactor import ReactorBase
    from pelephysics_amd.autotune import lookup
    e = lookup("drm19", "ReactorBDFAdaptiveMixed", "1d")
    assert e is not None
    m = models["drm19"]
    st = W.drm19_fresh(m, 512, seed=5)
    dev = torch.device("cuda:0")
    out = {}
    for tuned in (None, "1d"):
        R = ReactorBase.create("ReactorBDFAdaptiveMixed")
        R.init(2, st.ncells, mechanism="drm19", tuned=tuned, rtol=1e-10, atol=1e-16)
        if tuned:
            assert R.mechanism == e["kernels"]
        y = torch.from_numpy(st.y_1d()).to(dev)
        x = torch.from_numpy(st.rE.copy()).to(dev)
        R.react_1d(y, torch.zeros(st.ncells * m.n_species, dtype=torch.float64, device=dev), x,
                   torch.zeros(st.ncells, dtype=torch.float64, device=dev), 1e-5)
        out[tuned] = y.cpu().numpy().reshape(st.ncells, -1)
    a, b = out[None], out["1d"]
    Ya, Yb = a[:, :-1] / a[:, :-1].sum(1, keepdims=True), b[:, :-1] / b[:, :-1].sum(1, keepdims=True)
    assert np.all(np.abs(a[:, -1] - b[:, -1]) <= 1e-6 * np.abs(a[:, -1]) + 1e-10)
    assert np.all(np.abs(Ya - Yb) <= 1e-6 * np.abs(Ya) + 1e-10)
