// Wave-per-cell chemistry for the discrete adjoint (wave_adjoint.h): one
// 64-lane wavefront per cell, lane i < NEQ holding equation i (rhoY_i, then
// T) with its row of the Jacobian and of the LU factors in registers.
//
// Right-hand side: each species lane evaluates its NASA polynomials and
// exp(g/RT); concentrations and exponentials go to LDS; reactions are spread
// over the lanes (reaction j in lane j mod 64) and write their net rates of
// progress to LDS; each species lane then sums its own production rate over a
// compressed species->reaction table.  Temperature inversion, density and the
// energy equation use wave reductions.
//
// Jacobian: the same reaction-per-lane split with LDS double atomics
// (ds_add_f64) into an NEQ x NEQ matrix, then the temperature row, then each
// lane reads its row with the molecular-weight scaling of ReactorBDF.cpp:67-82.
//
// LU: row-per-lane Gaussian elimination with partial pivoting: the pivot is
// a wave max-reduction, the pivot row reaches the other lanes through
// v_readlane, and rows never move (each lane remembers the elimination step
// at which it was the pivot).
//
// (The batched adaptive integrators run several cells per wave, or one, in
// group_bdf.h.)
#pragma once
#include "chem_table.h"
#include "reactors.h"
#include "wave.h"

namespace ppc {

template <class T>
struct WaveChem {
  using TM = TabMech<T>;
  static constexpr int NS = T::NS, NR = T::NR, NEQ = T::NEQ;
  static_assert(NEQ <= 64, "one equation per lane");

  // per-lane constants of species k = lane (lanes >= NS hold species 0, masked)
  struct Lane {
    double tmid, mw, imw;
    int k;
    bool sp;  // lane is a species lane
  };
  template <typename LT>
  struct Shared {
    double nasa[NS][2][7];  // NASA7 coefficients (low, high range) per species
    double c[NS], a[NS], b[NS], q[NR];
    double g[NS], h[NS], wd[NS];
    double J[NEQ * NEQ];
    LT Jrow[NEQ][NEQ];  // Jacobian of the current Newton matrix, row i = lane i
  };

  template <class SH>
  PPC_DI static void load_lane(Lane& L, SH& sh, int lane) {
    const int k = lane < NS ? lane : 0;
    L.sp = lane < NS;
    L.k = k;
    for (int e = lane; e < NS * 7; e += 64) {
      sh.nasa[e / 7][0][e % 7] = T::nasaLo_(e);
      sh.nasa[e / 7][1][e % 7] = T::nasaHi_(e);
    }
    L.tmid = T::tmid_(k);
    L.mw = T::mw_(k);
    L.imw = T::imw_(k);
    __syncthreads();
  }
  template <class SH>
  PPC_DI static void coeffs(const Lane& L, const SH& sh, double Tv, double* a) {
    const int r = Tv < L.tmid ? 0 : 1;
#pragma unroll
    for (int i = 0; i < 7; ++i) a[i] = sh.nasa[L.k][r][i];
  }

  // mixture h (erg/g) and cp (erg/g/K), or e and cv, at Tv (Fuego.H CKHBMS/CKCPBS)
  template <bool ENTH, class SH>
  PPC_DI static void mix(const Lane& L, const SH& sh, double Y, double Tv, double& hm, double& cm) {
    double a[7];
    coeffs(L, sh, Tv, a);
    const double yw = L.sp ? Y * L.imw : 0.0;
    double hp = TM::h_poly(a, Tv, 1.0 / Tv), cp = TM::cp_poly(a, Tv);
    if (!ENTH) {
      hp -= 1.0;
      cp -= 1.0;
    }
    const double hs = wave::sum(yw * hp), cs = wave::sum(yw * cp);
    hm = hs * (RU * Tv);
    cm = cs * RU;
  }

  // GET_T_GIVEN_HY / GET_T_GIVEN_EY (Mechanisms/drm19/mechanism.H:2789-2890),
  // the Newton iteration of T_from_energy (chem_cell.h) with wave sums.
  template <bool ENTH, class SH>
  PPC_DI static double T_from_energy(const Lane& L, const SH& sh, double hv, double Y, double tguess) {
    const double tmin = 90.0, tmax = 4000.0;
    double hmin, cmin, hmax, cmax;
    mix<ENTH>(L, sh, Y, tmin, hmin, cmin);
    mix<ENTH>(L, sh, Y, tmax, hmax, cmax);
    if (hv < hmin) return tmin - (hmin - hv) / cmin;
    if (hv > hmax) return tmax - (hmax - hv) / cmax;
    double t1 = tguess;
    if (t1 < tmin || t1 > tmax) t1 = tmin + (tmax - tmin) / (hmax - hmin) * (hv - hmin);
    for (int i = 0; i < 200; ++i) {
      double h1, c1;
      mix<ENTH>(L, sh, Y, t1, h1, c1);
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
  template <class SH>
  PPC_DI static double T_from_energy_rt(const Lane& L, const SH& sh, int reactor_type, double hv, double Y,
                                        double tg) {
    return reactor_type == E_REACTOR ? T_from_energy<false>(L, sh, hv, Y, tg)
                                     : T_from_energy<true>(L, sh, hv, Y, tg);
  }

  // molar production rate (mol/m^3/s) of species `lane` from its concentration sc (mol/m^3)
  template <class SH>
  PPC_DI static double production_rate(const Lane& L, SH& sh, int lane, double sc, double Tv) {
    const double invT = 1.0 / Tv, logT = log(Tv);
    double cf[7];
    coeffs(L, sh, Tv, cf);
    if (L.sp) {
      const double eg = exp(TM::g_poly(cf, Tv, invT, logT));
      sh.a[lane] = 1.0 / eg;
      sh.b[lane] = sc * eg;
      sh.c[lane] = sc;
    }
    const double mixture = wave::sum(L.sp ? sc : 0.0);
    const double refC = 101325.0 / 8.31446 * invT, refCinv = 1.0 / refC;
    __syncthreads();
#pragma unroll 1
    for (int r0 = 0; r0 < NR; r0 += 64) {
      const int j = r0 + lane;
      if (j < NR) {
        double q;
        if (j < T::N_EL) {
          q = TM::template rate<0>(j, 0, sh.c, sh.a, sh.b, mixture, Tv, invT, logT, refC, refCinv);
        } else if (j < T::N_EL + T::N_TB) {
          q = TM::template rate<1>(j, 0, sh.c, sh.a, sh.b, mixture, Tv, invT, logT, refC, refCinv);
        } else if (j < T::N_EL + T::N_TB + T::N_TROE) {
          q = TM::template rate<2>(j, j - T::N_EL - T::N_TB, sh.c, sh.a, sh.b, mixture, Tv, invT, logT, refC,
                                   refCinv);
        } else {
          q = TM::template rate<3>(j, j - T::N_EL - T::N_TB, sh.c, sh.a, sh.b, mixture, Tv, invT, logT, refC,
                                   refCinv);
        }
        sh.q[j] = q;
      }
    }
    __syncthreads();
    double w = 0.0;
    if (L.sp) {
      const int e1 = T::spPtr_(lane + 1);
#pragma unroll 1
      for (int e = T::spPtr_(lane); e < e1; ++e) w += T::spNu_(e) * sh.q[T::spRx_(e)];
    }
    __syncthreads();
    return w;
  }

  // fKernelSpec (Source/Reactions/ReactorUtils.H:358-415) for the cell of this
  // wave: y = (rhoY_lane | T in lane NS), returns ydot of this lane.
  // Optional outputs (adjoint): the energy-inverted T, this lane's species
  // energy e_k or h_k (erg/g) and rho*cv or rho*cp.
  template <class SH>
  PPC_DI static double rhs(const Lane& L, SH& sh, int lane, double dt_save, int reactor_type, double y,
                           double rYsrc, double rhoe_init, double rhoesrc_ext, double* Tout = nullptr,
                           double* ekout = nullptr, double* rcout = nullptr) {
    const double rY = L.sp ? y : 0.0;
    const double rho = wave::sum(rY);
    const double rinv = 1.0 / rho;
    const double Y = rY * rinv;
    const double nrg = (rhoe_init + rhoesrc_ext * dt_save) * rinv;
    double Tv = wave::rdlane(y, NS);
    double cf[7];
    double ek, cmix, dummy;
    if (reactor_type == E_REACTOR) {
      Tv = T_from_energy<false>(L, sh, nrg, Y, Tv);
      coeffs(L, sh, Tv, cf);
      ek = TM::h_poly(cf, Tv, 1.0 / Tv) - 1.0;
      mix<false>(L, sh, Y, Tv, dummy, cmix);
    } else {
      Tv = T_from_energy<true>(L, sh, nrg, Y, Tv);
      coeffs(L, sh, Tv, cf);
      ek = TM::h_poly(cf, Tv, 1.0 / Tv);
      mix<true>(L, sh, Y, Tv, dummy, cmix);
    }
    const double RT = RU * Tv;
    ek = ek * RT * L.imw;  // CKUMS / CKHMS
    // RTY2WDOT with the CKWC unit conversion folded in (chem_cell.h wdot_mass)
    const double sc = (rho * Y * L.imw) * 1.0e6;
    double wd = production_rate(L, sh, lane, sc, Tv);
    wd = wd * 1.0e-6 * L.mw;
    const double v = L.sp ? wd + rYsrc : 0.0;
    const double src = rhoesrc_ext - wave::sum(v * ek);
    if (Tout) *Tout = Tv;
    if (ekout) *ekout = ek;
    if (rcout) *rcout = rho * cmix;
    return L.sp ? v : (lane == NS ? src * (rinv / cmix) : 0.0);
  }

  // The BDF Jacobian d f_i / d y_j (bdf_jacobian in chem_cell.h with s = 1)
  // into sh.Jrow (row i written by lane i).
  template <int CONSP, class SH>
  PPC_DI static void jacobian_rows(const Lane& L, SH& sh, int lane, double y) {
    constexpr int n = NEQ;
    const double sc = L.sp ? y * L.imw * 1.0e6 : 0.0;
    const double Tv = wave::rdlane(y, NS);
    const double invT = 1.0 / Tv, invT2 = invT * invT, logT = log(Tv);
    const double refC = 101325.0 / 8.31446 * invT, refCinv = 1.0 / refC;
    double cf[7];
    coeffs(L, sh, Tv, cf);
    const double hRT = TM::h_poly(cf, Tv, invT);
    const double cpR = TM::cp_poly(cf, Tv);
    const double dcR = cf[1] + Tv * (2.0 * cf[2] + Tv * (3.0 * cf[3] + Tv * 4.0 * cf[4]));
#pragma unroll 1
    for (int i = lane; i < n * n; i += 64) sh.J[i] = 0.0;
    if (L.sp) {
      sh.c[lane] = sc;
      sh.g[lane] = TM::g_poly(cf, Tv, invT, logT);
      sh.h[lane] = hRT;
      sh.wd[lane] = 0.0;
    }
    const double mixture = wave::sum(sc);
    __syncthreads();
#pragma unroll 1
    for (int r0 = 0; r0 < NR; r0 += 64) {
      const int j = r0 + lane;
      if (j < NR)
        TM::template jac_reaction<CONSP>(
            j, sh.c, sh.g, sh.h, mixture, Tv, invT, invT2, logT, refC, refCinv,
            [&](int idx, double v) { atomicAdd(&sh.J[idx], v); }, [&](int m, double v) { atomicAdd(&sh.wd[m], v); });
    }
    __syncthreads();
    // temperature row (mechanism.H:5098-5137)
    const double cR = (CONSP == 1) ? cpR : cpR - 1.0;
    const double eh = (CONSP == 1) ? hRT : hRT - 1.0;
    const double wdk = L.sp ? sh.wd[lane] : 0.0;
    const double cmix = wave::sum(L.sp ? cR * sc : 0.0);
    const double dcmixdT = wave::sum(L.sp ? dcR * sc : 0.0);
    const double ehmix = wave::sum(L.sp ? eh * wdk : 0.0);
    const double dehmixdT = wave::sum(L.sp ? invT * (cR - eh) * wdk + eh * sh.J[NS * n + lane] : 0.0);
    const double cmixinv = 1.0 / cmix;
    const double tmp1 = ehmix * cmixinv, tmp3 = cmixinv * Tv, tmp2 = tmp1 * tmp3;
    if (L.sp) {
      double d = 0.0;
#pragma unroll 1
      for (int mm = 0; mm < NS; ++mm) d += ((CONSP == 1) ? sh.h[mm] : sh.h[mm] - 1.0) * sh.J[lane * n + mm];
      sh.J[lane * n + NS] = tmp2 * cR - tmp3 * d;
    }
    if (lane == NS) sh.J[n * n - 1] = -tmp1 + tmp2 * dcmixdT - tmp3 * dehmixdT;
    __syncthreads();
    // DWDOT scaling (T column x1e-6, T row x1e6) and the transpose with
    // molecular-weight scaling: row i, column j = d f_i / d y_j
    const int i = lane < n ? lane : 0;
    const double wi = L.sp ? L.mw : 1.0;
    const double iwi = L.sp ? L.imw : 1.0;
    (void)iwi;
#pragma unroll
    for (int jj = 0; jj < n; ++jj) {
      double a;
      if (jj < NS) {
        const double Jv = (i < NS) ? sh.J[jj * n + i] : sh.J[jj * n + NS] * 1.0e6;
        a = Jv * wi * T::imw_(jj);
      } else {
        const double Jv = (i < NS) ? sh.J[NS * n + i] * 1.0e-6 : sh.J[n * n - 1];
        a = Jv * wi;
      }
      if (lane < n) sh.Jrow[lane][jj] = a;
    }
    __syncthreads();
  }
  template <class SH>
  PPC_DI static void jacobian_rows_rt(const Lane& L, SH& sh, int lane, int reactor_type, double y) {
    if (reactor_type == H_REACTOR)
      jacobian_rows<1>(L, sh, lane, y);
    else
      jacobian_rows<0>(L, sh, lane, y);
  }
};

// Row-per-lane LU with partial pivoting: lane i holds row i of the n x n
// matrix in a[0..n-1].  On return the lane that was the pivot at step k has
// kstep = k and its U row in a[k..n-1] (dinv = 1 / U_kk); the multipliers of
// every row sit in a[k] for the steps before its own; lane k holds pivlane =
// the pivot lane of step k.
template <int n, typename LT>
PPC_DI void wave_lu_factor(LT* a, int lane, int& kstep, int& pivlane, LT& dinv) {
  kstep = lane < n ? -1 : n;
  pivlane = 0;
  dinv = LT(0);
#pragma unroll
  for (int k = 0; k < n; ++k) {
    const bool cand = kstep < 0;
    const LT v = cand ? (LT)fabs(a[k]) : LT(-1);
    const LT m = wave::max(v);
    const int p = wave::first_lane(wave::ballot(cand && v == m), wave::first_lane(wave::ballot(cand), 0));
    const LT apk = wave::rdlane(a[k], p);
    const LT inv = LT(1) / apk;
    if (lane == p) {
      kstep = k;
      dinv = inv;
    }
    if (lane == k) pivlane = p;
    const bool elim = kstep < 0;
    const LT l = a[k] * inv;
    if (elim) a[k] = l;
#pragma unroll
    for (int j = k + 1; j < n; ++j) {
      const LT apj = wave::rdlane(a[j], p);
      if (elim) a[j] -= l * apj;
    }
  }
}

// Solve with the factors of wave_lu_factor; b = right-hand side entry of this
// lane's row, returns x_lane.
template <int n, typename LT>
PPC_DI double wave_lu_solve(const LT* a, int lane, int kstep, int pivlane, LT dinv, double b) {
#pragma unroll
  for (int k = 0; k < n; ++k) {
    const int p = wave::rdlane(pivlane, k);
    const double bp = wave::rdlane(b, p);
    if (kstep > k) b -= (double)a[k] * bp;
  }
  double x = 0.0;
#pragma unroll
  for (int k = n - 1; k >= 0; --k) {
    const int p = wave::rdlane(pivlane, k);
    const double xk = wave::rdlane(b * (double)dinv, p);
    if (lane == k) x = xk;
    if (kstep < k) b -= (double)a[k] * xk;
  }
  return x;
}

}  // namespace ppc
