"""Discrete adjoint of the fixed-step BDF (pp_chem_adjoint_1d) against central
finite differences through the oracle.

Workload: drm19 fresh CH4/air cells (T ~ U[1100, 1500] K, phi ~ U[0.6, 1.4],
1 atm) and partially burnt cells of the drm19 Gaussian-random-field grid,
integrated over 10 calls of ReactorBDF with dt = 1e-5 s (defaults: one
TRPZ substep per call), linear systems solved exactly.  Outputs T_final,
Y_CO, Y_OH; parameters ln A_j of all 84 reactions.

Criterion per cell and output: max_j |g_j - fd_j| <= 1e-4 max_j |fd_j| + noise,
with fd_j = (o(+delta) - o(-delta)) / (2 delta), delta = 1e-6, through the
oracle built with fused multiply-adds.  noise is the resolution of the
difference quotient itself: 4x the largest difference between the quotients
of the two oracle builds (with and without FMA contraction), plus 64 ulp(o) /
delta.  It matters where an output is at the rounding floor of the
integration (Y_CO ~ 1e-30 in a 612 K cell) or changes by less than its own
rounding at delta (T of a nearly inert cell); at least 10 (cell, output)
pairs must be resolved by the 1e-4 bound alone.
"""
import numpy as np
import pytest

import oracle
from pelephysics_amd import workloads as W

pytestmark = pytest.mark.gpu

DELTA = 1e-6
NMACRO, DT = 10, 1e-5


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch


def _cells(m):
    fresh = W.subsample(W.drm19_fresh(m, 4096), np.arange(0, 4096, 1024))
    g = W.drm19_grf(m, 4)
    warm = np.flatnonzero((g.T > 1400.0) & (g.T < 1900.0))[:4]
    grf = W.subsample(g, np.concatenate([[17], warm]))
    return W.CellState("drm19", np.concatenate([fresh.rY, grf.rY]), np.concatenate([fresh.T, grf.T]),
                       np.concatenate([fresh.rE, grf.rE]), np.concatenate([fresh.p, grf.p]))


def _oracle_outputs(O, st, dlnA, kCO, kOH):
    ns = st.rY.shape[1]
    O.set_dlnA(dlnA)
    y = st.y_1d()
    rX = st.rE.copy()
    p = oracle.bdf_params(ns, gmres_precond=-1)
    for _ in range(NMACRO):
        O.react_1d("bdf", y, np.zeros(st.ncells * ns), rX, np.zeros(st.ncells), DT, params=p)
    O.set_dlnA(None)
    y = y.reshape(-1, ns + 1)
    Y = y[:, :ns] / y[:, :ns].sum(1, keepdims=True)
    return np.stack([y[:, -1], Y[:, kCO], Y[:, kOH]], 1), y


def test_adjoint_matches_central_differences(models, oracles):
    torch = _torch()
    from pelephysics_amd.reactor import ReactorBase
    m = models["drm19"]
    O = oracles["drm19"]
    names = [s.name for s in m.species]
    kCO, kOH = names.index("CO"), names.index("OH")
    st = _cells(m)
    dev = torch.device("cuda:0")
    R = ReactorBase.create("ReactorBDF")
    R.init(2, st.ncells, mechanism="drm19")
    y = torch.from_numpy(st.y_1d().reshape(st.ncells, -1)).to(dev)
    rX = torch.from_numpy(st.rE.copy()).to(dev)
    out, grad = R.adjoint_1d(y, rX, NMACRO, DT)
    out, grad, yfin = out.cpu().numpy(), grad.cpu().numpy(), y.cpu().numpy()
    o0, yo = _oracle_outputs(O, st, None, kCO, kOH)
    # forward: the GPU runs the same map
    assert np.allclose(out, o0, rtol=1e-8, atol=1e-20)
    assert np.allclose(yfin, yo, rtol=1e-8, atol=1e-20)
    nr = m.n_reactions
    O2 = oracle.Mech(m, variant="nofma")
    fd = np.zeros((st.ncells, nr, 3))
    fd2 = np.zeros((st.ncells, nr, 3))
    for j in range(nr):
        e = np.zeros(nr)
        e[j] = DELTA
        fd[:, j] = (_oracle_outputs(O, st, e, kCO, kOH)[0] - _oracle_outputs(O, st, -e, kCO, kOH)[0]) / (2 * DELTA)
        fd2[:, j] = (_oracle_outputs(O2, st, e, kCO, kOH)[0] - _oracle_outputs(O2, st, -e, kCO, kOH)[0]) / (2 * DELTA)
    noise = 4 * np.abs(fd - fd2).max(1) + 64 * np.spacing(np.abs(o0)) / DELTA
    err = np.abs(grad - fd).max(1)
    tol = 1e-4 * np.abs(fd).max(1) + noise
    assert (err <= tol).all(), (err / np.maximum(np.abs(fd).max(1), 1e-300)).max(0)
    # the resolvable outputs are the ones that carry the check
    resolvable = np.abs(fd).max(1) * 1e-4 > noise
    assert resolvable.sum() >= 10
    # T_final of slowly reacting cells moves by less than its rounding at
    # delta = 1e-6: check it again at delta = 1e-3 (central differences,
    # truncation O(delta^2) relative)
    d3 = 1e-3
    fdT = np.zeros((st.ncells, nr))
    for j in range(nr):
        e = np.zeros(nr)
        e[j] = d3
        fdT[:, j] = (_oracle_outputs(O, st, e, kCO, kOH)[0][:, 0] - _oracle_outputs(O, st, -e, kCO, kOH)[0][:, 0]) / (2 * d3)
    errT = np.abs(grad[:, :, 0] - fdT).max(1)
    assert (errT <= 1e-4 * np.abs(fdT).max(1) + 64 * np.spacing(o0[:, 0]) / d3).all(), errT / np.abs(fdT).max(1)


def test_adjoint_autograd_wrapper(models):
    torch = _torch()
    from pelephysics_amd.reactor import ReactorBase
    from pelephysics_amd.adjoint import BDFSensitivity
    m = models["drm19"]
    st = W.subsample(W.drm19_grf(m, 4), np.array([5, 17, 40]))
    dev = torch.device("cuda:0")
    R = ReactorBase.create("ReactorBDF")
    R.init(2, st.ncells, mechanism="drm19")
    y0 = torch.from_numpy(st.y_1d().reshape(st.ncells, -1)).to(dev)
    rX = torch.from_numpy(st.rE.copy()).to(dev)
    dlnA = torch.zeros(m.n_reactions, dtype=torch.float64, device=dev, requires_grad=True)
    out = BDFSensitivity.apply(dlnA, y0, rX, R, NMACRO, DT)
    w = torch.tensor([1.0, 1e7, 1e8], dtype=torch.float64, device=dev)
    (out * w).sum().backward()
    _, grad = R.adjoint_1d(y0.clone(), rX, NMACRO, DT)
    ref = (grad * w[None, None, :]).sum(dim=(0, 2))
    assert torch.allclose(dlnA.grad, ref, rtol=1e-12, atol=0.0)
    with pytest.raises(ValueError):
        BDFSensitivity.apply(dlnA.detach() + 1.0, y0, rX, R, NMACRO, DT)


def test_adjoint_per_parameter_64_cells(models, oracles):
    """Every sensitivity on its own: 64 of the 4096 cells of the adjoint
    workload (every 64th), all 84 reactions and the three outputs, against
    central differences through the oracle at two steps.

    * delta = 1e-4: the quotient's truncation error is O(delta^2) ~ 1e-8
      relative and its rounding 100x below that at 1e-6; a sensitivity is
      resolved where 1e-4 |fd_j| exceeds the quotient's own noise (4x the FMA
      / non-FMA oracle spread plus 64 ulp(o) / delta), and at least 95% of
      the Y_CO and Y_OH sensitivities above 1e-3 of their cell's largest are;
    * delta = 1e-6: resolved where 1e-4 |fd_j| exceeds that quotient's
      measured error |fd(1e-6) - fd(1e-4)| plus its FMA spread and 64 ulp(o)
      / delta (at 1e-6 the
      quotients of outputs near 1e-12, such as Y_CO in cold cells, carry
      errors of a few 1e-4 from the integration's rounding).

    At both steps every resolved g_j is within 1e-3 |fd_j| + 1e-5 max_j |fd_j|
    (the adjoint's error is normwise, ~1e-6 of the largest sensitivity of the
    cell and output) and 99% of them within 1e-4 |fd_j|."""
    torch = _torch()
    from pelephysics_amd.reactor import ReactorBase
    m = models["drm19"]
    O = oracles["drm19"]
    names = [s.name for s in m.species]
    kCO, kOH = names.index("CO"), names.index("OH")
    st = W.subsample(W.drm19_fresh(m, 4096), np.arange(0, 4096, 64))
    dev = torch.device("cuda:0")
    R = ReactorBase.create("ReactorBDF")
    R.init(2, st.ncells, mechanism="drm19")
    y = torch.from_numpy(st.y_1d().reshape(st.ncells, -1)).to(dev)
    rX = torch.from_numpy(st.rE.copy()).to(dev)
    out, grad = R.adjoint_1d(y, rX, NMACRO, DT)
    grad = grad.cpu().numpy()
    o0, _ = _oracle_outputs(O, st, None, kCO, kOH)
    assert np.allclose(out.cpu().numpy(), o0, rtol=1e-8, atol=1e-20)
    nr = m.n_reactions
    O2 = oracle.Mech(m, variant="nofma")

    def quotients(Om, d):
        fd = np.zeros((st.ncells, nr, 3))
        for j in range(nr):
            e = np.zeros(nr)
            e[j] = d
            fd[:, j] = (_oracle_outputs(Om, st, e, kCO, kOH)[0] - _oracle_outputs(Om, st, -e, kCO, kOH)[0]) / (2 * d)
        return fd

    fd4, fd4b = quotients(O, 1e-4), quotients(O2, 1e-4)
    fd6, fd6b = quotients(O, DELTA), quotients(O2, DELTA)
    noise4 = 4 * np.abs(fd4 - fd4b) + (64 * np.spacing(np.abs(o0)) / 1e-4)[:, None, :]
    noise6 = np.abs(fd6 - fd4) + 4 * np.abs(fd6 - fd6b) + (64 * np.spacing(np.abs(o0)) / DELTA)[:, None, :]
    for fd, noise, tag in ((fd4, noise4, "1e-4"), (fd6, noise6, "1e-6")):
        resolved = 1e-4 * np.abs(fd) > noise
        err = np.abs(grad - fd)
        rel = np.where(resolved, err / np.where(resolved, np.abs(fd), 1.0), 0.0)
        within = (rel <= 1e-4)[resolved]
        # relative to the largest sensitivity of the same cell and output
        normwise = err / np.abs(fd).max(1, keepdims=True)
        print(f"delta {tag}: resolved {int(resolved.sum())} of {resolved.size} sensitivities, "
              f"{within.mean():.4%} within 1e-4, worst {float(rel.max()):.3g}; "
              f"worst normwise {float(normwise[np.abs(fd).max(1, keepdims=True).repeat(nr, 1) > 0].max()):.3g}")
        # every resolved sensitivity within 1e-3 |fd_j| + 1e-5 of the largest
        # sensitivity of its cell and output, 99% of them within 1e-4 |fd_j|:
        # the adjoint's error is ~1e-6 of that largest sensitivity (normwise),
        # which is up to 1e-3 of the sensitivities 1e-3 below it and more for
        # the smaller ones that an FD quotient happens to resolve
        floor = 1e-5 * np.abs(fd).max(1, keepdims=True)
        bad = resolved & (err > 1e-3 * np.abs(fd) + floor)
        assert not bad.any(), (tag, np.argwhere(bad)[:10])
        assert within.mean() >= 0.99, tag
        if tag == "1e-4":
            # Y_CO and Y_OH: nearly every sensitivity above 1e-3 of the cell's
            # largest is resolved (T_final of the nearly inert cells moves
            # less than its own rounding over the 1e-4 s)
            sig = np.abs(fd) >= 1e-3 * np.abs(fd).max(1, keepdims=True)
            assert (resolved & sig)[:, :, 1:].sum() >= 0.95 * sig[:, :, 1:].sum()
