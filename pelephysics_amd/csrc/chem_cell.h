// Per-cell chemistry on one lane: EOS wrappers, temperature inversion and the
// reactor right-hand side, templated on a generated mechanism struct M
// (pelephysics_amd/csrc/mech/<name>.h).
#pragma once
#include "ppc_common.h"

namespace ppc {

enum : int { E_REACTOR = 1, H_REACTOR = 2 };  // Source/Reactions/ReactorTypes.H:10-11

// rho = sum rhoY, Y = rhoY / rho (Source/Eos/Fuego.H:31-39 RY2RRinvY).
template <class M>
PPC_DI double rho_and_Y(const double* __restrict__ rY, double* __restrict__ Y, double& rinv) {
  double rho = 0.0;
#pragma unroll
  for (int k = 0; k < M::NS; ++k) rho += rY[k];
  rinv = 1.0 / rho;
#pragma unroll
  for (int k = 0; k < M::NS; ++k) Y[k] = rY[k] * rinv;
  return rho;
}

// GET_T_GIVEN_HY / GET_T_GIVEN_EY (Mechanisms/drm19/mechanism.H:2789-2890):
// Newton T <- T + (h - h(T))/cp(T), |dT| clamped to 100 K, stop (without
// applying the step) when |dT| < 1e-6 K, at most 200 iterations; linear
// extrapolation outside [90, 4000] K.
template <class M, bool ENTHALPY>
PPC_HEAVY double T_from_energy(const double hv, const double* __restrict__ Y, const double tguess) {
  const double tmin = 90.0, tmax = 4000.0;
  double hmin, hmax, c;
  if (ENTHALPY) {
    M::h_cp_mass(tmin, Y, hmin, c);
  } else {
    M::e_cv_mass(tmin, Y, hmin, c);
  }
  const double cmin = c;
  if (ENTHALPY) {
    M::h_cp_mass(tmax, Y, hmax, c);
  } else {
    M::e_cv_mass(tmax, Y, hmax, c);
  }
  const double cmax = c;
  if (hv < hmin) return tmin - (hmin - hv) / cmin;
  if (hv > hmax) return tmax - (hmax - hv) / cmax;
  double t1 = tguess;
  if (t1 < tmin || t1 > tmax) t1 = tmin + (tmax - tmin) / (hmax - hmin) * (hv - hmin);
  for (int i = 0; i < 200; ++i) {
    double h1, c1;
    if (ENTHALPY) {
      M::h_cp_mass(t1, Y, h1, c1);
    } else {
      M::e_cv_mass(t1, Y, h1, c1);
    }
    double dt = (hv - h1) / c1;
    if (dt > 100.0) {
      dt = 100.0;
    } else if (dt < -100.0) {
      dt = -100.0;
    } else if (fabs(dt) < 1.0e-6) {
      break;
    }
    t1 += dt;
  }
  return t1;
}

template <class M>
PPC_DI double T_from_energy_rt(int reactor_type, double hv, const double* Y, double tguess) {
  return reactor_type == E_REACTOR ? T_from_energy<M, false>(hv, Y, tguess)
                                   : T_from_energy<M, true>(hv, Y, tguess);
}

// RTY2WDOT (Source/Eos/Fuego.H:234-247): mass production rates in g/cm^3/s.
// The concentrations are formed directly in SI (mol/m^3) so the CKWC
// conversion (x1e6 in, x1e-6 out, mechanism.H:5671-5687) is folded in.
template <class M>
PPC_DI void wdot_mass(const double rho, const double T, const double* __restrict__ Y,
                      double* __restrict__ wdot) {
  double sc[M::NS];
#pragma unroll
  for (int k = 0; k < M::NS; ++k) sc[k] = (rho * Y[k] * M::imw(k)) * 1.0e6;
  M::production_rate(wdot, sc, T);
#pragma unroll
  for (int k = 0; k < M::NS; ++k) wdot[k] = wdot[k] * 1.0e-6 * M::mw(k);
}

// fKernelSpec (Source/Reactions/ReactorUtils.H:358-415) for one cell:
// T from the energy (guess = y[NS]), species rates plus external sources, and
// the temperature equation with cv (e-form) or cp (h-form).
template <class M>
PPC_HEAVY void f_kernel_spec(const double dt_save, const int reactor_type, const double* __restrict__ y,
                          double* __restrict__ ydot, const double rhoe_init, const double rhoesrc_ext,
                          const double* __restrict__ rYsrc) {
  double Y[M::NS];
  double rinv;
  const double rho = rho_and_Y<M>(y, Y, rinv);
  const double nrg = (rhoe_init + rhoesrc_ext * dt_save) * rinv;
  double T = y[M::NS];
  double ei[M::NS];
  double cmix, dummy;
  if (reactor_type == E_REACTOR) {
    T = T_from_energy<M, false>(nrg, Y, T);
    M::e_RT(ei, T);
    M::e_cv_mass(T, Y, dummy, cmix);
  } else {
    T = T_from_energy<M, true>(nrg, Y, T);
    M::h_RT(ei, T);
    M::h_cp_mass(T, Y, dummy, cmix);
  }
  const double RT = RU * T;
#pragma unroll
  for (int k = 0; k < M::NS; ++k) ei[k] = ei[k] * RT * M::imw(k);  // CKUMS / CKHMS
  double wd[M::NS];
  wdot_mass<M>(rho, T, Y, wd);
  double src = rhoesrc_ext;
#pragma unroll
  for (int k = 0; k < M::NS; ++k) {
    const double v = wd[k] + rYsrc[k];
    ydot[k] = v;
    src -= v * ei[k];
  }
  ydot[M::NS] = src * (rinv / cmix);
}

// The BDF Jacobian in (rhoY, T) variables, row-major A[i*n+j] = s * d f_i / d y_j
// (ReactorBDF.cpp:67-82: RTY2JAC -> DWDOT at the integrated T, species block
// scaled by mw_i/mw_j, T column by mw_i, T row by 1/mw_j).  Jc is a
// scratch array of n*n doubles; A may alias Jc.  JM is double* or a
// LaneArr<double> (lane_ws.h).
template <class M, class JM>
PPC_HEAVY void bdf_jacobian(const int reactor_type, const double* __restrict__ y, const double s,
                         const JM Jc, const JM A) {
  constexpr int NS = M::NS, n = NS + 1;
  double sc[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) sc[k] = y[k] * M::imw(k) * 1.0e6;  // CKYTCR + DWDOT (x1e6)
  if (reactor_type == H_REACTOR) {
    M::template jacobian<1>(Jc, sc, y[NS]);
  } else {
    M::template jacobian<0>(Jc, sc, y[NS]);
  }
  // DWDOT scaling of the T column (x1e-6) and T row (x1e6)
  for (int k = 0; k < NS; ++k) {
    Jc[NS * n + k] *= 1.0e-6;
    Jc[k * n + NS] *= 1.0e6;
  }
  // transpose (column-major -> row-major) with molecular-weight scaling, in
  // place; chunks of 8 pairs load before they store (Jc and A may alias)
  constexpr int U = 8;
  for (int i = 0; i < n; ++i) {
    const double wi = (i < NS) ? M::mw_rt(i) : 1.0;
    const double iwi = (i < NS) ? M::imw_rt(i) : 1.0;
    A[i * n + i] = s * Jc[i * n + i];
    for (int j0 = i + 1; j0 < n; j0 += U) {
      double cji[U], cij[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (j0 + u < n) {
          cji[u] = Jc[(j0 + u) * n + i];
          cij[u] = Jc[i * n + j0 + u];
        }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (j0 + u < n) {
          const int j = j0 + u;
          const double wj = (j < NS) ? M::mw_rt(j) : 1.0;
          const double iwj = (j < NS) ? M::imw_rt(j) : 1.0;
          const double a_ij = cji[u] * wi * iwj;  // d f_i / d y_j
          const double a_ji = cij[u] * wj * iwi;  // d f_j / d y_i
          A[i * n + j] = s * a_ij;
          A[j * n + i] = s * a_ji;
        }
    }
  }
}

}  // namespace ppc
