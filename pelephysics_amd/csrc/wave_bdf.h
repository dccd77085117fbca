// Batched adaptive BDF with one 64-lane wavefront per cell.
//
// Lane i < NEQ holds equation i of the cell (rhoY_i, then T): its entries of
// the BDF difference array, the iterate, the error weights, row i of the
// Jacobian and row i of the LU factors of the Newton matrix, all in
// registers.  The integrator's control flow (step size, order, Newton
// convergence, error test) depends only on wave-wide reductions, so it is
// uniform: a wave never executes a branch that its cell does not take.
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
// Newton matrix: S^-1 (I - c J) S with S the error weights, factored in LU_T
// (float: FP32 LU, FP64 residuals and rates) by row-per-lane Gaussian
// elimination with partial pivoting: the pivot is a wave max-reduction, the
// pivot row reaches the other lanes through v_readlane, and rows never move
// (each lane remembers the elimination step at which it was the pivot).
//
// The step/order/Newton logic is the NDF family of adaptive_cell (reactors.h),
// itself the algorithm of SciPy's BDF integrator.
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

PPC_DI double wave_rms(double x, double sc, bool act, int n) {
  const double v = act ? x / sc : 0.0;
  return sqrt(wave::sum(v * v) / n);
}

// D[0..order] <- (R U)^T D[0..order] for a step-size ratio `factor` (change_D
// in reactors.h).  Lane 6i+j forms RU[i][j] = sum_k R[i][k] U[k][j] with
// R[i][k] = prod_{m=1..i} (m-1-factor k)/m and U = R at factor 1, in the
// order of the sequential form; the entries reach every lane by v_readlane.
PPC_DI void wave_change_D(double* D, const int order, const double factor, const int lane) {
  const int i = lane / (AMAXORD + 1), j = lane % (AMAXORD + 1);
  double ru = 0.0;
  if (i <= order && j <= order) {
#pragma unroll 1
    for (int k = 0; k <= order; ++k) {
      double r = 1.0, u = 1.0;
#pragma unroll 1
      for (int m = 1; m <= i; ++m) r *= (m - 1 - factor * k) / (double)m;
#pragma unroll 1
      for (int m = 1; m <= k; ++m) u *= (m - 1 - 1.0 * j) / (double)m;
      ru += r * u;
    }
  }
  double tmp[AMAXORD + 1];
#pragma unroll
  for (int c = 0; c <= AMAXORD; ++c) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k <= AMAXORD; ++k)
      if (k <= order) s += wave::rdlane(ru, k * (AMAXORD + 1) + c) * D[k];
    tmp[c] = s;
  }
#pragma unroll
  for (int c = 0; c <= AMAXORD; ++c)
    if (c <= order) D[c] = tmp[c];
}

enum : int { WPH_F0 = 0, WPH_F1 = 1, WPH_NEWTON = 2 };

// Adaptive BDF for the cell of this wave from t = 0 to t_bound; y in/out
// (this lane's equation).  Returns the accepted step count or a negative
// status.  Written as a loop with one right-hand side per trip (initial-step
// probes, then Newton iterations) so that the RHS, the Jacobian, the LU and
// change_D each have a single call site in the kernel.
template <class T, typename LT>
PPC_DI int wave_cell(const AdaptParamsDev& ap, const int reactor_type, const typename WaveChem<T>::Lane& L,
                     typename WaveChem<T>::template Shared<LT>& sh, const int lane, double& y, const double src,
                     const double rhoe, const double rhoesrc, const double t_bound, int* cnt) {
  using W = WaveChem<T>;
  constexpr int n = W::NEQ;
  const bool act = lane < n;
  const double rtol = ap.rtol, atol = ap.atol;
  const double newton_tol = fmax(10 * 2.220446049250313e-16 / rtol, fmin(0.03, sqrt(rtol)));
  double D[AMAXORD + 3];
#pragma unroll
  for (int k = 0; k < AMAXORD + 3; ++k) D[k] = 0.0;
  LT LU[n];
#pragma unroll
  for (int k = 0; k < n; ++k) LU[k] = LT(0);
  int kstep = 0, pivlane = 0;
  LT dinv = LT(0);
  const double y0 = act ? y : 0.0;
  D[0] = y0;
  double ypred = 0, scale = 1, psi = 0, yv = y0, d = 0, f0 = 0, lus = 1;
  int nsteps = 0, nrhs = 0, njac = 0, nlu = 0, nnewton = 0, nrej = 0, status = 0;
  double t = 0.0, t_new = 0.0, t_eval = 0.0, h_abs = 0.0, h0 = 0.0, c = 0.0, dy_norm_old = -1.0;
  double min_step = 0.0, safety = 0.9;
  int order = 1, n_equal = 0, k_newton = 0, phase = WPH_F0;
  bool lu_valid = false, current_jac = false;
  // every path through the loop either accepts a step (bounded by max_steps)
  // or shrinks h until the minimum-step test fails; the RHS budget is a
  // second bound against non-finite step sizes
  const long long rhs_cap = 64LL * ap.max_steps + 4096;
  for (;;) {
    if (nrhs >= rhs_cap) {
      status = -2;
      break;
    }
    const double f = W::rhs(L, sh, lane, t_eval, reactor_type, yv, src, rhoe, rhoesrc);
    nrhs++;
    // what to prepare before the next right-hand side
    bool jac = false, jac_pred = false, new_step = false, new_attempt = false, retry = false, cd = false;
    double cd_fac = 1.0;
    if (phase == WPH_F0) {
      f0 = f;
      if (ap.first_step > 0) {
        h_abs = ap.first_step;
        jac = new_step = true;
      } else {
        // initial step (Hairer, Norsett & Wanner I, II.4)
        scale = atol + fabs(y0) * rtol;
        const double d0 = wave_rms(y0, scale, act, n), d1 = wave_rms(f, scale, act, n);
        h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
        h0 = fmin(h0, t_bound);
        yv = y0 + h0 * f;
        t_eval = t + h0;
        phase = WPH_F1;
        continue;
      }
    } else if (phase == WPH_F1) {
      const double d1 = wave_rms(f0, scale, act, n);
      const double d2 = wave_rms(f - f0, scale, act, n) / h0;
      const double h1 = (d1 <= 1e-15 && d2 <= 1e-15) ? fmax(1e-6, h0 * 1e-3) : sqrt(0.01 / fmax(d1, d2));
      h_abs = fmin(fmin(100 * h0, h1), t_bound);
      jac = new_step = true;
    } else {
      // one modified-Newton iteration with f = f(t_new, yv)
      nnewton++;
      const bool finite = wave::ballot(act && !isfinite(f)) == 0ull;
      bool failed = !finite, converged = false;
      if (finite) {
        double dy = (c * f - psi - d) / lus;
        dy = wave_lu_solve<n, LT>(LU, lane, kstep, pivlane, dinv, act ? dy : 0.0);
        dy *= lus;
        const double dy_norm = wave_rms(dy, scale, act, n);
        const double rate = dy_norm_old < 0 ? -1.0 : dy_norm / dy_norm_old;
        if (rate >= 0 && (rate >= 1 || pow(rate, 4 - k_newton) / (1 - rate) * dy_norm > newton_tol)) {
          failed = true;
        } else {
          yv += dy;
          d += dy;
          if (dy_norm == 0 || (rate >= 0 && rate / (1 - rate) * dy_norm < newton_tol)) {
            converged = true;
          } else {
            dy_norm_old = dy_norm;
            k_newton++;
            if (k_newton == 4) failed = true;
          }
        }
      }
      const int n_iter = k_newton + 1;
      if (failed) {
        if (!current_jac) {
          // refresh the Jacobian at the predictor and retry the same step
          jac = jac_pred = retry = true;
          current_jac = true;
          lu_valid = false;
        } else {
          h_abs *= 0.5;
          cd = true;
          cd_fac = 0.5;
          n_equal = 0;
          lu_valid = false;
          nrej++;
          new_attempt = true;
        }
      } else if (converged) {
        safety = 0.9 * 9.0 / (double)(8 + n_iter);
        scale = atol + rtol * fabs(yv);
        const double error_norm = wave_rms(ndf_econst(order) * d, scale, act, n);
        if (error_norm > 1) {
          const double fac = fmax(0.2, safety * pow(error_norm, -1.0 / (order + 1)));
          h_abs *= fac;
          cd = true;
          cd_fac = fac;
          n_equal = 0;
          nrej++;
          new_attempt = true;
        } else {
          // accept: D[order+2] = d - D[order+1]; D[order+1] = d; D[k] += D[k+1], k = order..0
          nsteps++;
          n_equal++;
          t = t_new;
          double Dn[AMAXORD + 3];
#pragma unroll
          for (int k = 0; k < AMAXORD + 3; ++k) Dn[k] = D[k];
#pragma unroll
          for (int k = 1; k < AMAXORD + 3; ++k) {
            if (k == order + 2) Dn[k] = d - D[k - 1];
            if (k == order + 1) Dn[k] = d;
          }
#pragma unroll
          for (int k = AMAXORD + 1; k >= 0; --k)
            if (k <= order) Dn[k] += Dn[k + 1];
#pragma unroll
          for (int k = 0; k < AMAXORD + 3; ++k) D[k] = Dn[k];
          if (n_equal >= order + 1) {
            double em = INFINITY, ep = INFINITY, Dm = 0.0, Dp = 0.0;
#pragma unroll
            for (int k = 0; k < AMAXORD + 3; ++k) {
              if (k == order) Dm = D[k];
              if (k == order + 2) Dp = D[k];
            }
            if (order > 1) em = wave_rms(ndf_econst(order - 1) * Dm, scale, act, n);
            if (order < AMAXORD) ep = wave_rms(ndf_econst(order + 1) * Dp, scale, act, n);
            const double norms[3] = {em, error_norm, ep};
            double best = -1.0;
            int delta = 0;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              const double fct = (norms[i] == 0.0) ? INFINITY : pow(norms[i], -1.0 / (order + i));
              if (fct > best) {
                best = fct;
                delta = i - 1;
              }
            }
            order += delta;
            const double fac = fmin(10.0, safety * best);
            h_abs *= fac;
            cd = true;
            cd_fac = fac;
            n_equal = 0;
            lu_valid = false;
          }
          if (!(t < t_bound)) break;
          new_step = true;
        }
      }
    }
    // ---- prepare the next right-hand side
    if (jac) {
      W::jacobian_rows_rt(L, sh, lane, reactor_type, jac_pred ? ypred : D[0]);
      njac++;
    }
    if (phase != WPH_NEWTON) {
      // first step: D[1] = h f0
      D[1] = f0 * h_abs;
    }
    // the change_D calls of one transition, in order: the step-size change of
    // the decision, the minimum-step clamp at a new step, the clamp to t_bound
    for (int stage = 0; stage < 3; ++stage) {
      bool doit = false;
      double fac = 1.0;
      if (stage == 0) {
        doit = cd;
        fac = cd_fac;
      } else if (stage == 1) {
        if (new_step) {
          if (nsteps >= ap.max_steps) {
            status = -2;
            break;
          }
          min_step = 10 * fabs(nextafter(t, INFINITY) - t);
          if (h_abs < min_step) {
            doit = true;
            fac = min_step / h_abs;
            h_abs = min_step;
            n_equal = 0;
          }
          current_jac = false;
        }
      } else if (new_step || new_attempt) {
        if (h_abs < min_step) {
          status = -3;
          break;
        }
        t_new = t + h_abs;
        if (t_new - t_bound > 0) {
          t_new = t_bound;
          doit = true;
          fac = fabs(t_new - t) / h_abs;
          n_equal = 0;
          lu_valid = false;
        }
      }
      if (doit) wave_change_D(D, order, fac, lane);
    }
    if (status != 0) break;
    if (new_step || new_attempt) {
      const double h = t_new - t;
      h_abs = fabs(h);
      const double inv_alpha = 1.0 / ndf_alpha(order);
      double sum = 0.0, ps = 0.0;
#pragma unroll
      for (int k = 0; k <= AMAXORD; ++k)
        if (k <= order) sum += D[k];
#pragma unroll
      for (int k = 1; k <= AMAXORD; ++k)
        if (k <= order) ps += D[k] * bdf_gamma(k);
      ypred = sum;
      scale = atol + rtol * fabs(sum);
      psi = ps * inv_alpha;
      c = h * inv_alpha;
    }
    if (new_step || new_attempt || retry) {
      if (!lu_valid) {
        // S^-1 (I - c J) S with S the current error weights
        lus = scale;
        const double si = 1.0 / lus;
#pragma unroll
        for (int j = 0; j < n; ++j) {
          const double sj = wave::rdlane(lus, j);
          const double m = (lane == j ? 1.0 : 0.0) - c * (act ? (double)sh.Jrow[lane][j] : 0.0);
          LU[j] = (LT)(si * m * sj);
        }
        wave_lu_factor<n, LT>(LU, lane, kstep, pivlane, dinv);
        nlu++;
        lu_valid = true;
      }
      yv = ypred;
      d = 0.0;
      dy_norm_old = -1.0;
      k_newton = 0;
    }
    t_eval = t_new;
    phase = WPH_NEWTON;
  }
  y = D[0];
  cnt[0] = nrhs;
  cnt[1] = njac;
  cnt[2] = nlu;
  cnt[3] = nnewton;
  cnt[4] = nrej;
  return status == 0 ? nsteps : status;
}

// Cell input/output of the 1-D (YCOrder) form, one wave per cell.
struct WaveIO1D {
  double* y;
  const double* ysrc;
  double* rX;
  const double* rXsrc;
  template <class W, class SH>
  PPC_DI void load(int c, int lane, const typename W::Lane& L, const SH&, int, double& y0, double& src,
                   double& rhoe, double& rhoesrc) const {
    y0 = lane < W::NEQ ? y[(long long)c * W::NEQ + lane] : 0.0;
    src = lane < W::NS ? ysrc[(long long)c * W::NS + lane] : 0.0;
    rhoe = rX[c];
    rhoesrc = rXsrc[c];
  }
  template <class W, class SH>
  PPC_DI void store(int c, int lane, const typename W::Lane& L, const SH&, int, double yv, double rhoe,
                    double rhoesrc, double dt, int) const {
    if (lane < W::NEQ) y[(long long)c * W::NEQ + lane] = yv;
    if (lane == 0) rX[c] = rhoe + dt * rhoesrc;
  }
};

// Cell input/output of the box (Array4) form with the initial and final
// temperature inversions of ReactorBDF.cpp:376-389,461-478.
struct WaveIOBox {
  BoxArgs b;
  PPC_DI void ijk(int c, int& i, int& j, int& k) const {
    i = b.lo[0] + c % b.len[0];
    j = b.lo[1] + (c / b.len[0]) % b.len[1];
    k = b.lo[2] + c / (b.len[0] * b.len[1]);
  }
  template <class W, class SH>
  PPC_DI void load(int c, int lane, const typename W::Lane& L, const SH& sh, int reactor_type, double& y0,
                   double& src, double& rhoe, double& rhoesrc) const {
    int i, j, k;
    ijk(c, i, j, k);
    const double rY = L.sp ? b.rY(i, j, k, lane) : 0.0;
    src = L.sp ? b.rYsrc(i, j, k, lane) : 0.0;
    const double rho = wave::sum(rY);
    const double Y = rY * (1.0 / rho);
    rhoe = b.rE(i, j, k, 0);
    rhoesrc = b.rEsrc(i, j, k, 0);
    const double T0 = W::T_from_energy_rt(L, sh, reactor_type, rhoe / rho, Y, b.T(i, j, k, 0));
    y0 = L.sp ? rY : (lane == W::NS ? T0 : 0.0);
  }
  template <class W, class SH>
  PPC_DI void store(int c, int lane, const typename W::Lane& L, const SH& sh, int reactor_type, double yv,
                    double rhoe, double rhoesrc, double dt, int nsteps) const {
    int i, j, k;
    ijk(c, i, j, k);
    const double rY = L.sp ? yv : 0.0;
    const double rho = wave::sum(rY);
    const double Y = rY * (1.0 / rho);
    if (L.sp) b.rY(i, j, k, lane) = yv;
    const double rE_new = rhoe + dt * rhoesrc;
    const double Tn = W::T_from_energy_rt(L, sh, reactor_type, rE_new / rho, Y, wave::rdlane(yv, W::NS));
    if (lane == 0) {
      b.rE(i, j, k, 0) = rE_new;
      b.T(i, j, k, 0) = Tn;
      if (b.FC.p) b.FC(i, j, k, 0) = (double)nsteps;
    }
  }
};

// Waves per SIMD the register allocation is held to.
#ifndef PPC_WAVE_OCC
#define PPC_WAVE_OCC __attribute__((amdgpu_waves_per_eu(2)))
#endif

// Persistent kernel: each wave takes cells from the queue until it is empty.
template <class T, typename LT, class IO>
__global__ void __launch_bounds__(64) PPC_WAVE_OCC wave_bdf_kernel(ReactParamsDev P, IO io, double dt, QueueArgs q,
                                                      CellStats st) {
  using W = WaveChem<T>;
  __shared__ typename W::template Shared<LT> sh;
  const int lane = wave::lane_id();
  typename W::Lane L;
  W::load_lane(L, sh, lane);
  for (;;) {
    int cell = 0;
    if (lane == 0) cell = atomicAdd(q.counter, 1);
    cell = wave::rdlane(cell, 0);
    if (cell >= q.ncells) break;
    const int cid = q.order ? q.order[cell] : cell;
    double y, src, rhoe, rhoesrc;
    io.template load<W>(cid, lane, L, sh, P.reactor_type, y, src, rhoe, rhoesrc);
    int cnt[5];
    const int r = wave_cell<T, LT>(P.adapt, P.reactor_type, L, sh, lane, y, src, rhoe, rhoesrc, dt, cnt);
    io.template store<W>(cid, lane, L, sh, P.reactor_type, y, rhoe, rhoesrc, dt, r);
    if (lane == 0) store_counters(st, cid, r, cnt);
  }
}

}  // namespace ppc
