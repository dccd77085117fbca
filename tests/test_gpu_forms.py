"""Forms and reactor types not covered elsewhere: the ReactorRK64 box form
(ReactorRK64.cpp:166-320: T from energy first, exact step counts in FC) and
the constant-volume e-reactor (reactor_type 1, rho e = rho h - p) for the
adaptive, batched (one cell per lane and a group of lanes per cell) and
FP64-LU batched integrators, against the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle
from pelephysics_amd import workloads as W
from test_gpu_parity import _torch, _box_tensors, _compare_y, _oracle_1d, _run_1d, _close

pytestmark = pytest.mark.gpu


def test_rk64_box_form_ragged_halo(models, oracles):
    torch = _torch()
    from pelephysics_amd.reactor import ReactorBase, Array4, library
    m = models["LiDryer"]
    st = W.subsample(W.lidryer_h2air(m, 200), np.arange(105))
    nx, ny, nz = 5, 3, 7
    comp = _box_tensors(torch, st, m, (nx, ny, nz))
    rng = np.random.default_rng(4)
    ysrc = 1.0 * rng.standard_normal(st.rY.shape) * st.rY
    ysrc -= ysrc.sum(1, keepdims=True) / m.n_species
    esrc = 1.0 * rng.standard_normal(st.ncells) * np.abs(st.rE)
    rY, rYs, T = comp(st.rY, m.n_species), comp(ysrc, m.n_species), comp(st.T[:, None] + 2.0, 1)
    rE, rEs, FC = comp(st.rE[:, None], 1), comp(esrc[:, None], 1), comp(np.zeros((105, 1)), 1)

    def a4(t):
        a = Array4()
        a.p = t.data_ptr()
        a.nstride, a.kstride, a.jstride = t.stride()[0], t.stride()[1], t.stride()[2]
        a.begin[:] = [-1, -1, -1]
        a.end[:] = [nx + 1, ny + 1, nz + 1]
        a.ncomp = t.shape[0]
        return a

    R = ReactorBase.create("ReactorRK64")
    R.init(2, 105, mechanism="LiDryer")
    lo, hi = (C.c_int * 3)(0, 0, 0), (C.c_int * 3)(nx - 1, ny - 1, nz - 1)
    views = [a4(t) for t in (rY, rYs, T, rE, rEs, FC)]
    dt, tm = C.c_double(1e-5), C.c_double(0.0)
    rc = library().pp_chem_react_box(R._h, lo, hi, *[C.byref(v) for v in views], None, C.byref(dt), C.byref(tm),
                                     None)
    R._check(rc)
    torch.cuda.synchronize()
    inner = (slice(None), slice(1, 1 + nz), slice(1, 1 + ny), slice(1, 1 + nx))
    gY = rY[inner].cpu().numpy().reshape(m.n_species, -1).T
    gT = T[inner].cpu().numpy().reshape(-1)
    gE = rE[inner].cpu().numpy().reshape(-1)
    oY = np.ascontiguousarray(st.rY.T).copy()
    oT, orE = st.T.copy() + 2.0, st.rE.copy()
    oret, oFC = oracles["LiDryer"].react_box("rk64", oY, np.ascontiguousarray(ysrc.T).copy(), oT, orE, esrc, 1e-5)
    assert rc == oret
    assert (FC[inner].cpu().numpy().reshape(-1) == oFC).all()
    _close("T", gT, oT)
    _close("Y", gY / gY.sum(1, keepdims=True), oY.T / oY.T.sum(1, keepdims=True))
    _close("rE", gE, orE)


@pytest.mark.parametrize("kernels,ident,ode", [
    ("drm19", "ReactorBDFAdaptive", {}),
    ("drm19", "ReactorBDFAdaptiveMixed", {}),
    ("drm19_table", "ReactorBDFAdaptiveMixed", {"batched_cells_per_wave": 4}),
    ("drm19", "ReactorBDFBatchedFP64", {}),
    ("drm19_table", "ReactorBDFBatchedFP64", {"batched_cells_per_wave": 2}),
])
def test_e_reactor_adaptive_forms(models, oracles, kernels, ident, ode):
    """Constant-volume cells (rho e) with sources, against the oracle's
    adaptive BDF at rtol 1e-13 / atol 1e-20."""
    torch = _torch()
    m = models["drm19"]
    st = W.drm19_fresh(m, 96, seed=9)
    st.rE[:] = st.rE - st.p  # rho e = rho h - p
    rng = np.random.default_rng(5)
    ysrc = 1e-2 * rng.standard_normal((st.ncells, m.n_species)) * st.rY
    ysrc -= ysrc.sum(1, keepdims=True) / m.n_species
    esrc = 1e-2 * rng.standard_normal(st.ncells) * np.abs(st.rE)
    ret, y, rX, R = _run_1d(torch, kernels, ident, st, ysrc, esrc, 1e-5, reactor_type=1, rtol=1e-10, atol=1e-16,
                            **ode)
    assert R.stats()["nfailed"] == 0
    _, oy, orX, _ = _oracle_1d(oracles["drm19"], "adaptive", st, ysrc, esrc, 1e-5, reactor_type=1,
                               params=oracle.adapt_params(rtol=1e-13, atol=1e-20))
    _compare_y(m.n_species + 1, y, oy)
    _close("rX", rX, orX)
