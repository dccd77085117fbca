"""The reference's ignition-delay known-answer test on the GPU: the dodecane_lu
homogeneous reactor of Testing/Exec/IgnitionDelay (T0 = 600 K, p = 5 atm,
h-form, rtol = atol = 1e-10) integrated by the batched mixed-precision BDF in
device calls of dt / ndt, the delay measured as computeIgnitionDelay.py does;
it must be within 10 call lengths of 0.0763 s (check_ignDelay.py:16-32), as
the oracle's is (tests/test_oracle_ignition.py)."""
import numpy as np
import pytest

from pelephysics_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ode", [{}, {"batched_cells_per_wave": 1}])
def test_dodecane_ignition_delay_kat_gpu(models, ode):
    import torch
    from pelephysics_amd.reactor import ReactorBase
    assert torch.cuda.is_available()
    m = models["dodecane_lu"]
    st0 = W.ignition_delay_state(m)
    R = ReactorBase.create("ReactorBDFAdaptiveMixed")
    R.init(2, 1, mechanism="dodecane_lu", rtol=1e-10, atol=1e-10, **ode)
    dev = torch.device("cuda:0")
    ys = torch.zeros(m.n_species, dtype=torch.float64, device=dev)
    xs = torch.zeros(1, dtype=torch.float64, device=dev)
    state = {}

    def advance(start, dtc, k):
        if start:
            state["y"] = torch.from_numpy(st0.y_1d()).to(dev)
            state["x"] = torch.from_numpy(st0.rE.copy()).to(dev)
        Ts = torch.empty(k, dtype=torch.float64, device=dev)
        for i in range(k):
            R.react_1d(state["y"], ys, state["x"], xs, dtc)
            Ts[i] = state["y"][m.n_species]
        return Ts.cpu().numpy()

    t_ign, unc = W.ignition_delay(advance, 600.0)
    print(f"ignition delay {t_ign:.5g} s +/- {unc:.3g} s (reference 0.0763 s) {ode}")
    assert abs(t_ign - 0.0763) <= 10.0 * unc
