"""Synthetic cell states for the benchmark and test workloads.

Every workload returns the reactor state in the reference's CGS units as
numpy arrays: ``rY`` (ncells, N) in g/cm^3, ``T`` (ncells,) in K, ``rE``
(ncells,) = rho*h in erg/cm^3 (h-form, reactor_type 2) and zero sources.

* ``lidryer_h2air``: H2/air, T ~ U[1000,1400] K, p = 1 atm, phi ~ U[0.5,1.5],
  radicals (H, O, OH, HO2, H2O2) at Y = 1e-8, H2O = 0 (32^3 cells by default).
* ``drm19_grf``: CH4/air; mixture fraction from a periodic Gaussian random
  field (correlation length 8 cells) mapped through the normal CDF, progress
  c ~ U[0,1] interpolating linearly between fresh gas at 300 K and complete
  combustion products at 2200 K (64^3 cells by default).
* ``drm19_fresh``: CH4/air, T ~ U[1100,1500] K, phi ~ U[0.6,1.4], p = 1 atm.
* ``dodecane_uneven``: n-C12H26/air cells of very uneven stiffness, shuffled:
  70% cold fresh gas (T ~ U[300,700] K, phi = 1, p = 1 atm), 25% near
  ignition (T ~ U[900,1200] K, phi ~ U[0.7,1.3], p ~ U[1,40] atm), 5% complete
  combustion products at 2300 K (2^20 cells by default).
* ``heptane_ignition``: n-C7H16/air near ignition, T ~ U[800,1300] K,
  phi ~ U[0.5,2.0], p ~ U[10,50] atm (131072 cells by default).
* ``premixed``: a small fuel/air state for any compiled mechanism (tests).

Seeds are fixed (numpy PCG64) so every run sees the same cells.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .mechgen.model import Mechanism, RU, PATM

SEED = 20261202


@dataclass
class CellState:
    mech: str
    rY: np.ndarray  # (ncells, N)
    T: np.ndarray  # (ncells,)
    rE: np.ndarray  # (ncells,)
    p: np.ndarray  # (ncells,)

    @property
    def ncells(self) -> int:
        return self.T.shape[0]

    def y_1d(self) -> np.ndarray:
        """YCOrder vector y[icell*(N+1)+k], k = N is T."""
        return np.ascontiguousarray(np.concatenate([self.rY, self.T[:, None]], axis=1).reshape(-1))


def _nasa(m: Mechanism):
    lo = np.array([s.low for s in m.species])
    hi = np.array([s.high for s in m.species])
    tmid = np.array([s.tmid for s in m.species])
    return lo, hi, tmid


def h_mass(m: Mechanism, T: np.ndarray, Y: np.ndarray) -> np.ndarray:
    """Mixture enthalpy (erg/g) for T (ncells,), Y (ncells, N), NASA7."""
    lo, hi, tmid = _nasa(m)
    T = np.asarray(T, float)[:, None]
    a = np.where(T < tmid[None, :], lo[None, :, :].transpose(2, 0, 1), hi[None, :, :].transpose(2, 0, 1))
    hrt = a[0] + a[1] / 2 * T + a[2] / 3 * T ** 2 + a[3] / 4 * T ** 3 + a[4] / 5 * T ** 4 + a[5] / T
    imw = np.array([s.imw for s in m.species])
    return (Y * imw[None, :] * hrt).sum(axis=1) * RU * T[:, 0]


def mean_mw(m: Mechanism, Y: np.ndarray) -> np.ndarray:
    imw = np.array([s.imw for s in m.species])
    return 1.0 / (Y * imw[None, :]).sum(axis=1)


def _finish(m: Mechanism, name: str, Y: np.ndarray, T: np.ndarray, p: np.ndarray) -> CellState:
    Y = Y / Y.sum(axis=1, keepdims=True)
    rho = p * mean_mw(m, Y) / (RU * T)
    rY = rho[:, None] * Y
    rE = rho * h_mass(m, T, Y)
    return CellState(name, np.ascontiguousarray(rY), np.ascontiguousarray(T),
                     np.ascontiguousarray(rE), p)


def _mass_from_moles(m: Mechanism, X: dict, ncells: int) -> np.ndarray:
    mw = np.array([s.mw for s in m.species])
    Y = np.zeros((ncells, m.n_species))
    for name, x in X.items():
        Y[:, m.species_index(name)] = x * mw[m.species_index(name)]
    return Y / Y.sum(axis=1, keepdims=True)


def lidryer_h2air(m: Mechanism, ncells: int = 32 ** 3, seed: int = SEED) -> CellState:
    rng = np.random.Generator(np.random.PCG64(seed))
    T = rng.uniform(1000.0, 1400.0, ncells)
    phi = rng.uniform(0.5, 1.5, ncells)
    Y = _mass_from_moles(m, {"H2": 2.0 * phi, "O2": np.ones(ncells), "N2": 3.76 * np.ones(ncells)},
                         ncells)
    for r in ("H", "O", "OH", "HO2", "H2O2"):
        Y[:, m.species_index(r)] = 1e-8
    Y[:, m.species_index("H2O")] = 0.0
    return _finish(m, m.name, Y, T, np.full(ncells, PATM))


def _gaussian_field(n: int, corr: float, rng) -> np.ndarray:
    """Periodic Gaussian random field on n^3, Gaussian filter sigma = corr cells, unit variance."""
    w = rng.standard_normal((n, n, n))
    k = np.fft.fftfreq(n)
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    filt = np.exp(-0.5 * (2 * np.pi * corr) ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    f = np.real(np.fft.ifftn(np.fft.fftn(w) * filt))
    return (f - f.mean()) / f.std()


def _normal_cdf(x: np.ndarray) -> np.ndarray:
    erf = np.vectorize(math.erf)
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0)))


def drm19_grf(m: Mechanism, n: int = 64, seed: int = SEED) -> CellState:
    """Mixture-fraction field x progress variable, cell order i fastest (x, y, z).

    The correlation length is 8 cells on the 64^3 grid and scales with n on
    smaller grids (n/8 cells)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Z = _normal_cdf(_gaussian_field(n, 8.0 * n / 64.0, rng)).transpose(2, 1, 0).reshape(-1)  # z-major
    ncells = Z.shape[0]
    c = rng.uniform(0.0, 1.0, ncells)
    mw = {s.name: s.mw for s in m.species}
    iCH4, iO2, iN2 = (m.species_index(s) for s in ("CH4", "O2", "N2"))
    iCO2, iH2O = m.species_index("CO2"), m.species_index("H2O")
    Yf = np.zeros((ncells, m.n_species))
    Yf[:, iCH4] = Z
    Yf[:, iO2] = 0.233 * (1 - Z)
    Yf[:, iN2] = 0.767 * (1 - Z)
    # complete combustion CH4 + 2 O2 -> CO2 + 2 H2O of the same mixture
    nF = Z / mw["CH4"]
    nO = 0.233 * (1 - Z) / mw["O2"]
    burnt = np.minimum(nF, nO / 2.0)
    Yb = Yf.copy()
    Yb[:, iCH4] = (nF - burnt) * mw["CH4"]
    Yb[:, iO2] = (nO - 2 * burnt) * mw["O2"]
    Yb[:, iCO2] = burnt * mw["CO2"]
    Yb[:, iH2O] = 2 * burnt * mw["H2O"]
    Y = (1 - c)[:, None] * Yf + c[:, None] * Yb
    T = (1 - c) * 300.0 + c * 2200.0
    return _finish(m, m.name, Y, T, np.full(ncells, PATM))


def drm19_fresh(m: Mechanism, ncells: int = 4096, seed: int = SEED,
                Trange=(1100.0, 1500.0), phirange=(0.6, 1.4)) -> CellState:
    rng = np.random.Generator(np.random.PCG64(seed))
    T = rng.uniform(*Trange, ncells)
    phi = rng.uniform(*phirange, ncells)
    Y = _mass_from_moles(m, {"CH4": phi, "O2": 2.0 * np.ones(ncells),
                             "N2": 2.0 * 3.76 * np.ones(ncells)}, ncells)
    return _finish(m, m.name, Y, T, np.full(ncells, PATM))


def subsample(st: CellState, idx: np.ndarray) -> CellState:
    return CellState(st.mech, np.ascontiguousarray(st.rY[idx]), np.ascontiguousarray(st.T[idx]),
                     np.ascontiguousarray(st.rE[idx]), st.p[idx])


# Fuel of each mechanism and its stoichiometric O2 per fuel molecule.
FUELS = {"LiDryer": ("H2", 0.5), "drm19": ("CH4", 2.0), "dodecane_lu": ("NC12H26", 18.5),
         "heptane_lu_88sk": ("NC7H16", 11.0), "CH4_lean_qss": ("CH4", 2.0),
         "dodecane_lu_qss": ("NC12H26", 18.5), "heptane_lu_qss": ("NC7H16", 11.0),
         "LuEthylene_qss": ("C2H4", 3.0)}


def _fuel_air_Y(m: Mechanism, phi: np.ndarray) -> np.ndarray:
    fuel, s = FUELS[m.name]
    n = phi.shape[0]
    return _mass_from_moles(m, {fuel: phi, "O2": s * np.ones(n), "N2": 3.76 * s * np.ones(n)}, n)


def _products_Y(m: Mechanism, phi: np.ndarray) -> np.ndarray:
    """Complete combustion products (CO2, H2O, N2 and the excess O2 or fuel)."""
    fuel, s = FUELS[m.name]
    comp = m.species[m.species_index(fuel)].composition
    nC, nH = comp.get("C", 0.0), comp.get("H", 0.0)
    n = phi.shape[0]
    burnt = np.minimum(phi, 1.0)  # moles of fuel burnt per s moles of O2
    X = {fuel: phi - burnt, "O2": s * (1.0 - burnt), "N2": 3.76 * s * np.ones(n),
         "H2O": 0.5 * nH * burnt}
    if nC > 0:
        X["CO2"] = nC * burnt
    return _mass_from_moles(m, X, n)


def dodecane_uneven(m: Mechanism, ncells: int = 2 ** 20, seed: int = SEED) -> CellState:
    rng = np.random.Generator(np.random.PCG64(seed))
    n_fresh, n_ign = int(0.70 * ncells), int(0.25 * ncells)
    n_burnt = ncells - n_fresh - n_ign
    T = np.concatenate([rng.uniform(300.0, 700.0, n_fresh), rng.uniform(900.0, 1200.0, n_ign),
                        np.full(n_burnt, 2300.0)])
    phi = np.concatenate([np.ones(n_fresh), rng.uniform(0.7, 1.3, n_ign), np.ones(n_burnt)])
    p = np.concatenate([np.full(n_fresh, PATM), rng.uniform(1.0, 40.0, n_ign) * PATM, np.full(n_burnt, PATM)])
    Y = np.concatenate([_fuel_air_Y(m, phi[:n_fresh + n_ign]), _products_Y(m, phi[n_fresh + n_ign:])])
    perm = rng.permutation(ncells)
    return _finish(m, m.name, Y[perm], T[perm], p[perm])


def heptane_ignition(m: Mechanism, ncells: int = 131072, seed: int = SEED) -> CellState:
    rng = np.random.Generator(np.random.PCG64(seed))
    T = rng.uniform(800.0, 1300.0, ncells)
    phi = rng.uniform(0.5, 2.0, ncells)
    p = rng.uniform(10.0, 50.0, ncells) * PATM
    return _finish(m, m.name, _fuel_air_Y(m, phi), T, p)


def premixed(m: Mechanism, ncells: int, seed: int = SEED) -> CellState:
    """Small fuel/air state of any compiled mechanism: LiDryer and drm19 as
    their workloads above, the larger mechanisms near ignition."""
    if m.name == "LiDryer":
        return lidryer_h2air(m, ncells, seed)
    if m.name == "drm19":
        return drm19_fresh(m, ncells, seed)
    if m.name in ("dodecane_lu", "dodecane_lu_qss", "CH4_lean_qss", "LuEthylene_qss"):
        rng = np.random.Generator(np.random.PCG64(seed))
        T = rng.uniform(900.0, 1200.0, ncells)
        phi = rng.uniform(0.7, 1.3, ncells)
        p = rng.uniform(1.0, 40.0, ncells) * PATM
        return _finish(m, m.name, _fuel_air_Y(m, phi), T, p)
    return heptane_ignition(m, ncells, seed)


def ignition_delay_state(m: Mechanism, T0: float = 600.0, p: float = 5066250.0,
                         equiv_ratio: float = 7.928571428571429, ncells: int = 1) -> CellState:
    """The homogeneous-reactor start of the reference's ignition-delay test
    (Testing/Exec/IgnitionDelay/GPU_misc.H:59-65 with inputs.0d_firstpass):
    X_O2 = 1 / (phi / 18.5 + 1 + 0.79 / 0.21), X_N2 = 0.79 X_O2 / 0.21, the
    fuel the rest; T0 = 600 K, p = 5 atm, h-form."""
    fuel, _ = FUELS[m.name]
    xo2 = 1.0 / (equiv_ratio / 18.5 + 1.0 + 0.79 / 0.21)
    xn2 = 0.79 * xo2 / 0.21
    Y = _mass_from_moles(m, {"O2": xo2 * np.ones(ncells), "N2": xn2 * np.ones(ncells),
                             fuel: (1.0 - xo2 - xn2) * np.ones(ncells)}, ncells)
    return _finish(m, m.name, Y, np.full(ncells, T0), np.full(ncells, p))


def ignition_delay(advance, T0: float, dt: float = 2.0, ndt: int = 2000):
    """The ignition delay as the reference's test measures it
    (Testing/Exec/IgnitionDelay/computeIgnitionDelay.py, exec_ignDelay.sh):
    calls of dt / ndt from the start until T exceeds T0 + 400 K, the delay
    being (dt / ndt) x that call count; first with dt = 2 s, ndt = 2000, then
    again with dt = 1.1 x the first estimate and ndt = 4000.  `advance(start,
    dt_call, ncalls)` integrates ncalls calls of dt_call (from the initial
    state when start is true, else from where the previous calls ended) and
    returns the temperature after each.  Returns (delay, uncertainty dt / ndt
    of the refined pass)."""
    def first_crossing(dtc, n):
        done, chunk = 0, 50
        while done < n:
            k = min(chunk, n - done)
            Ts = advance(done == 0, dtc, k)
            hit = np.argwhere(np.asarray(Ts) > T0 + 400.0)
            if hit.size:
                return done + int(hit[0, 0]) + 1
            done += k
        raise RuntimeError("no ignition within the calls")
    est = dt / ndt * first_crossing(dt / ndt, ndt)
    dt2, ndt2 = est * 1.1, 2 * ndt
    return dt2 / ndt2 * first_crossing(dt2 / ndt2, ndt2), dt2 / ndt2
