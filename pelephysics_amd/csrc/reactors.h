// Per-cell integrators (one lane = one cell) and the launch kernels.
//
//  * bdf_parity_cell: ReactorBDF's fixed-step BDF1/TRPZ/BDF2/BDF3 with one or
//    more linearised Newton iterations and left-preconditioned GMRES
//    (Source/Reactions/ReactorBDF.cpp:187-285, ReactorBDFsolver.H:12-499);
//  * rk64_cell: ReactorRK64's 6-stage low-storage RK with the absolute-error,
//    never-reject controller (ReactorRK64.cpp:73-144);
//  * adaptive_cell: variable-order (1-5) backward-difference integrator with
//    error control (Shampine-Reichelt quasi-constant step NDF family), modified
//    Newton with a dense FP64 LU and Jacobian reuse.
#pragma once
#include "chem_cell.h"
#include "erk.h"
#include "lane_ws.h"
#include "reactors_fwd.h"

namespace ppc {

// ------------------------------------------------------ BDF (parity) ----
// BDFParams (ReactorBDF.H:20-31): the reference's 4-decimal truncated rows.
PPC_DI double tcoeff(int s, int m) {
  const double t[4][4] = {{1.0, -1.0, 0.0, 0.0},
                          {1.0, -1.0, 0.0, 0.0},
                          {1.0, -1.3333, 0.3333, 0.0},
                          {1.0, -1.6363, 0.8181, -0.1818}};
  return t[s][m];
}
PPC_DI double fcoeff(int s, int m) {
  const double f[4][4] = {{1.0, 0.0, 0.0, 0.0},
                          {0.5, 0.5, 0.0, 0.0},
                          {0.6666, 0.0, 0.0, 0.0},
                          {0.5454, 0.0, 0.0, 0.0}};
  return f[s][m];
}

// Preconditioner solve M^-1 X (ReactorBDFsolver.H:12-91): 0 none, 1 Jacobi,
// 2 symmetric Gauss-Seidel (D+L) D^-1 (D+U).
template <int n, class AM>
PPC_DI void precond(double* __restrict__ out, const AM A, const double* __restrict__ X, int type) {
  if (type == 0) {
    for (int i = 0; i < n; ++i) out[i] = X[i];
  } else if (type == 1) {
    for (int i = 0; i < n; ++i) out[i] = X[i] / A[i * n + i];
  } else {
    double yd[n];
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < i; ++j) s += A[i * n + j] * yd[j];
      yd[i] = (X[i] - s) / A[i * n + i];
    }
    for (int i = 0; i < n; ++i) yd[i] = A[i * n + i] * yd[i];
    for (int i = n - 1; i >= 0; --i) {
      double s = 0.0;
      for (int j = n - 1; j > i; --j) s += A[i * n + j] * out[j];
      out[i] = (yd[i] - s) / A[i * n + i];
    }
  }
}

// Restarted GMRES with warm start (ReactorBDFsolver.H:384-499): solves A x = b
// for the left-preconditioned system; x0 is the initial guess (overwritten).
// V ((n+1) x n Krylov basis) and H ((n+1) x n Hessenberg matrix) are
// caller-provided storage.  Returns false when a NaN/Inf appears in H.
template <int n, class AM, class VM>
PPC_DI bool gmres(const AM A, const VM V, const VM H, const double* __restrict__ b, double* __restrict__ x0,
                  double* __restrict__ x, int ptype, int restarts, int kspiters, double tol) {
  constexpr int K = n;
  double cs[K], sn[K], be[K + 1], yv[K + 1];
  double r[n], t1[n], t2[n];
  for (int i = 0; i < (K + 1) * K; ++i) H[i] = 0.0;
  for (int i = 0; i < K; ++i) { cs[i] = 0.0; sn[i] = 0.0; }
  for (int i = 0; i <= K; ++i) { be[i] = 0.0; yv[i] = 0.0; }
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += A[i * n + j] * x0[j];
    t1[i] = b[i] - s;
  }
  precond<n>(r, A, t1, ptype);
  for (int i = 0; i < n; ++i) x[i] = x0[i];
  double residnorm = 0.0;
  bool lucky = false;
  for (int it = 0; it < restarts; ++it) {
    for (int i = 0; i < n; ++i) x0[i] = x[i];
    double beta = 0.0;
    for (int i = 0; i < n; ++i) beta += r[i] * r[i];
    beta = sqrt(beta);
    for (int i = 0; i < n; ++i) V[i] = r[i] / beta;
    be[0] = beta;
    int opt = K;
    for (int k = 0; k < kspiters; ++k) {
      for (int i = 0; i < n; ++i) x[i] = x0[i];
      for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += A[i * n + j] * V[k * n + j];
        t1[i] = s;
      }
      precond<n>(t2, A, t1, ptype);
      for (int i = 0; i <= k; ++i) {
        double s = 0.0;
        for (int l = 0; l < n; ++l) s += t2[l] * V[i * n + l];
        H[i * K + k] = s;
      }
      for (int i = 0; i <= k; ++i)
        for (int l = 0; l < n; ++l) t2[l] = t2[l] - H[i * K + k] * V[i * n + l];
      double hn = 0.0;
      for (int l = 0; l < n; ++l) hn += t2[l] * t2[l];
      hn = sqrt(hn);
      H[(k + 1) * K + k] = hn;
      lucky = false;
      if (hn > tol * 0.01) {
        for (int l = 0; l < n; ++l) V[(k + 1) * n + l] = t2[l] / hn;
      } else {
        if (!isfinite(hn)) return false;
        lucky = true;
      }
      for (int i = 0; i <= k - 1; ++i) {
        const double tmp = cs[i] * H[i * K + k] + sn[i] * H[(i + 1) * K + k];
        H[(i + 1) * K + k] = -sn[i] * H[i * K + k] + cs[i] * H[(i + 1) * K + k];
        H[i * K + k] = tmp;
      }
      const double rr = H[k * K + k], ss = H[(k + 1) * K + k];
      double c, sg;
      if (rr == 0) {
        c = 0.0;
        sg = 1.0;
      } else {
        const double d = sqrt(rr * rr + ss * ss);
        c = rr / d;
        sg = ss / d;
      }
      H[k * K + k] = c * H[k * K + k] + sg * H[(k + 1) * K + k];
      H[(k + 1) * K + k] = 0.0;
      cs[k] = c;
      sn[k] = sg;
      be[k + 1] = -sg * be[k];
      be[k] = c * be[k];
      residnorm = fabs(be[k + 1]);
      if (lucky || residnorm <= tol) {
        opt = k + 1;
        break;
      }
    }
    // uppertrisolve (ReactorBDFsolver.H:368-382): the column index runs up to
    // opt inclusive; column K lies outside the Hessenberg matrix and is zero.
    for (int i = opt - 1; i >= 0; --i) {
      double v = be[i];
      for (int j = i + 1; j <= opt; ++j) v -= (j < K ? H[i * K + j] : 0.0) * yv[j];
      yv[i] = v / H[i * K + i];
    }
    for (int i = 0; i < opt; ++i)
      for (int l = 0; l < n; ++l) x[l] = x[l] + yv[i] * V[i * n + l];
    if (residnorm <= tol || lucky) break;
  }
  return true;
}

// Fixed-step BDF core; y (n) in/out, returns nsubsteps or -1 on NaN.
template <class M>
PPC_DI int bdf_parity_cell(const BDFParamsDev& p, const int reactor_type, double* __restrict__ y,
                           const double rhoe_init, const double rhoesrc, const double* __restrict__ rYsrc,
                           const double dt_react, const LaneWS& w) {
  constexpr int n = M::NEQ;
  constexpr bool WS = uses_ws<M>();
  double soln[n], soln_n[n], nm1[n], nm2[n], ds[n], ds0[n], rhs[n], f[n], fn[n];
  double A_p[WS ? 1 : n * n], V_p[WS ? 1 : (n + 1) * n], H_p[WS ? 1 : (n + 1) * n];
  const auto A = lane_array<WS>(A_p, w, 0);
  const auto V = lane_array<WS>(V_p, w, n * n);
  const auto H = lane_array<WS>(H_p, w, n * n + (n + 1) * n);
  const double dt = dt_react / (double)p.nsubsteps;
  const double dt_inv = 1.0 / dt;
  double current_time = 0.0;
  for (int i = 0; i < n; ++i) {
    soln[i] = soln_n[i] = nm1[i] = nm2[i] = y[i];
    ds[i] = ds0[i] = 0.0;
  }
  const int first = (p.scheme >= 2) ? 1 : p.scheme;
  const int change = p.scheme - 2;
  for (int st = 0; st < p.nsubsteps; ++st) {
    const int ts = (st > change) ? p.scheme : first;
    for (int it = 0; it < p.nonlinear_iters; ++it) {
      for (int i = 0; i < n; ++i) ds0[i] = ds[i];
      // get_bdf_matrix_and_rhs (ReactorBDF.cpp:36-110)
      bdf_jacobian<M>(reactor_type, soln, -fcoeff(ts, 0), A, A);
      f_kernel_spec<M>(current_time, reactor_type, soln, f, rhoe_init, rhoesrc, rYsrc);
      if (ts == 1) {
        f_kernel_spec<M>(current_time, reactor_type, soln_n, fn, rhoe_init, rhoesrc, rYsrc);
      } else {
        for (int i = 0; i < n; ++i) fn[i] = 0.0;
      }
      const double t0 = tcoeff(ts, 0), t1 = tcoeff(ts, 1), t2 = tcoeff(ts, 2), t3 = tcoeff(ts, 3);
      const double f0 = fcoeff(ts, 0), f1 = fcoeff(ts, 1);
      for (int i = 0; i < n; ++i) {
        A[i * n + i] += t0 * dt_inv;
        double r = -t0 * soln[i] * dt_inv;
        r += -t1 * soln_n[i] * dt_inv;
        r += -t2 * nm1[i] * dt_inv;
        r += -t3 * nm2[i] * dt_inv;
        r += f0 * f[i];
        r += f1 * fn[i];
        rhs[i] = r;
      }
      if (!gmres<n>(A, V, H, rhs, ds0, ds, p.gmres_precond, p.gmres_restarts, p.gmres_kspiters, p.gmres_tol))
        return -1;
      for (int i = 0; i < n; ++i) soln[i] += ds[i];
      double nrm = 0.0;
      for (int i = 0; i < n; ++i) nrm += rhs[i] * rhs[i];
      if (sqrt(nrm) <= p.nonlin_tol) break;
    }
    for (int i = 0; i < n; ++i) {
      nm2[i] = nm1[i];
      nm1[i] = soln_n[i];
      soln_n[i] = soln[i];
    }
    current_time += dt;
  }
  for (int i = 0; i < n; ++i) y[i] = soln_n[i];
  return p.nsubsteps;
}

// ---------------------------------------------------------- RK64 --------
// RK64Params (ReactorRK64.H:9-40)
PPC_DI double rk_alpha(int s) {
  const double a[6] = {0.218150805229859, 0.256702469801519, 0.527402592007520,
                       0.0484864267224467, 1.24517071533530, 0.412366034843237};
  return a[s];
}
PPC_DI double rk_beta(int s) {
  const double b[6] = {-0.113554138044166, -0.215118587818400, -0.0510152146250577,
                       -1.07992686223881, -0.248664241213447, 0.0};
  return b[s];
}
PPC_DI double rk_err(int s) {
  const double e[6] = {-0.0554699315064507, 0.158481845574980, -0.0905918835751907,
                       -0.219084567203338, 0.164022338959433, 0.0426421977505659};
  return e[s];
}

// The stepping of ReactorRK64::react on n equations; rhs(t, y, ydot) is the
// right-hand side at time t from the start of the call.
template <int n, class RHS>
PPC_DI int rk64_core(const RK64ParamsDev& p, double* __restrict__ y, const double dt_react, RHS&& rhs) {
  double carry[n], err[n], yd[n];
  const double time_out = dt_react;
  double current_time = 0.0;
  for (int i = 0; i < n; ++i) carry[i] = y[i];
  double dt_rk = dt_react / (double)p.nsub_guess;
  const double dt_min = dt_react / (double)p.nsub_max;
  const double dt_max = dt_react / (double)p.nsub_min;
  int nsteps = 0;
  while (current_time < time_out) {
    for (int i = 0; i < n; ++i) err[i] = 0.0;
    for (int st = 0; st < 6; ++st) {
      rhs(current_time, y, yd);
      const double a = rk_alpha(st), b = rk_beta(st), e = rk_err(st);
      for (int i = 0; i < n; ++i) {
        err[i] += e * dt_rk * yd[i];
        y[i] = carry[i] + a * dt_rk * yd[i];
        carry[i] = y[i] + b * dt_rk * yd[i];
      }
    }
    current_time += dt_rk;
    nsteps++;
    double max_err = 1e-50;
    for (int i = 0; i < n; ++i) max_err = fabs(err[i]) > max_err ? fabs(err[i]) : max_err;
    if (max_err < p.atol) {
      dt_rk = fmin(dt_max, dt_rk * pow(p.atol / max_err, 0.25));
    } else {
      dt_rk = fmax(dt_min, dt_rk * pow(p.atol / max_err, 0.2));
    }
    dt_rk = fmin(dt_rk, time_out - current_time);
  }
  return nsteps;
}

template <class M>
PPC_DI int rk64_cell(const RK64ParamsDev& p, const int reactor_type, double* __restrict__ y,
                     const double rhoe_init, const double rhoesrc, const double* __restrict__ rYsrc,
                     const double dt_react) {
  return rk64_core<M::NEQ>(p, y, dt_react, [&](double t, const double* yy, double* yd) {
    f_kernel_spec<M>(t, reactor_type, yy, yd, rhoe_init, rhoesrc, rYsrc);
  });
}

// ------------------------------------------------ adaptive BDF (FP64) ---
// The algorithm follows SciPy's scipy/integrate/_ivp/bdf.py (BSD-3-Clause,
// Copyright (c) 2001-2002 Enthought, Inc. 2003-2024, SciPy Developers; the
// notice and license are in THIRD_PARTY_NOTICES.md at the repository root).
constexpr int AMAXORD = 5;

PPC_DI double ndf_kappa(int k) {
  const double kap[6] = {0.0, -0.1850, -1.0 / 9.0, -0.0823, -0.0415, 0.0};
  return kap[k];
}
PPC_DI double bdf_gamma(int k) {
  const double g[6] = {0.0, 1.0, 1.5, 1.8333333333333333, 2.083333333333333, 2.283333333333333};
  return g[k];
}
PPC_DI double ndf_alpha(int k) { return (1.0 - ndf_kappa(k)) * bdf_gamma(k); }
PPC_DI double ndf_econst(int k) {
  return k > AMAXORD ? 0.0 : ndf_kappa(k) * bdf_gamma(k) + 1.0 / (k + 1);
}

template <int n>
PPC_DI double rms_norm(const double* __restrict__ x, const double* __restrict__ sc) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    const double v = x[i] / sc[i];
    s += v * v;
  }
  return sqrt(s / n);
}

// In-place LU with partial pivoting (row-major); a is a T* or a LaneArr<T>.
// The compiler cannot tell that the rows read and the row written differ, so
// an element-by-element update would wait for each store before the next
// load; the updates therefore go in chunks of LU_CHUNK columns whose loads are
// all issued before their stores, with a chunk of the pivot row held in
// registers across the rows below it.  Each element receives the same
// operations in the same order as in the textbook k-i-j loop.
#ifndef PPC_LU_CHUNK
#define PPC_LU_CHUNK 8
#endif
constexpr int LU_CHUNK = PPC_LU_CHUNK;

template <int n, typename T, class AM>
PPC_HEAVY void lu_factor(const AM a, int* __restrict__ piv) {
  constexpr int U = LU_CHUNK;
  for (int k = 0; k < n; ++k) {
    int p = k;
    T mx = fabs(a[k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const T v = fabs(a[i * n + k]);
      if (v > mx) {
        mx = v;
        p = i;
      }
    }
    piv[k] = p;
    if (p != k)
      for (int j0 = 0; j0 < n; j0 += U) {
        T rk[U], rp[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (j0 + u < n) {
            rk[u] = a[k * n + j0 + u];
            rp[u] = a[p * n + j0 + u];
          }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (j0 + u < n) {
            a[k * n + j0 + u] = rp[u];
            a[p * n + j0 + u] = rk[u];
          }
      }
    const T inv = T(1) / a[k * n + k];
    for (int i0 = k + 1; i0 < n; i0 += U) {  // multipliers l_ik = a_ik / a_kk
      T l[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (i0 + u < n) l[u] = a[(i0 + u) * n + k];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (i0 + u < n) a[(i0 + u) * n + k] = l[u] * inv;
    }
    for (int j0 = k + 1; j0 < n; j0 += U) {
      T rk[U];
#pragma unroll
      for (int u = 0; u < U; ++u) rk[u] = (j0 + u < n) ? a[k * n + j0 + u] : T(0);
      for (int i = k + 1; i < n; ++i) {
        const T l = a[i * n + k];
        T ri[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (j0 + u < n) ri[u] = a[i * n + j0 + u];
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (j0 + u < n) a[i * n + j0 + u] = ri[u] - l * rk[u];
      }
    }
  }
}

template <int n, typename T, class AM>
PPC_HEAVY void lu_solve(const AM a, const int* __restrict__ piv, double* __restrict__ b) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) {
      const double t = b[k];
      b[k] = b[piv[k]];
      b[piv[k]] = t;
    }
  for (int i = 0; i < n; ++i) {
    double s = b[i];
    for (int j = 0; j < i; ++j) s -= (double)a[i * n + j] * b[j];
    b[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= (double)a[i * n + j] * b[j];
    b[i] = s / (double)a[i * n + i];
  }
}

// D[0..order] <- (R U)^T D[0..order] for a step-size ratio `factor`.
template <int n>
PPC_HEAVY void change_D(double* __restrict__ D, const int order, const double factor) {
  double R[AMAXORD + 1][AMAXORD + 1], U[AMAXORD + 1][AMAXORD + 1];
  for (int pass = 0; pass < 2; ++pass) {
    const double f = pass == 0 ? factor : 1.0;
    for (int j = 0; j <= order; ++j) {
      double c = 1.0;
      for (int i = 0; i <= order; ++i) {
        const double m = (i == 0) ? 1.0 : (j == 0 ? 0.0 : (i - 1 - f * j) / (double)i);
        c *= m;
        if (pass == 0) R[i][j] = c; else U[i][j] = c;
      }
    }
  }
  double RU[AMAXORD + 1][AMAXORD + 1];
  for (int i = 0; i <= order; ++i)
    for (int j = 0; j <= order; ++j) {
      double s = 0.0;
      for (int k = 0; k <= order; ++k) s += R[i][k] * U[k][j];
      RU[i][j] = s;
    }
  for (int c = 0; c < n; ++c) {
    double tmp[AMAXORD + 1];
    for (int i = 0; i <= order; ++i) {
      double s = 0.0;
      for (int k = 0; k <= order; ++k) s += RU[k][i] * D[k * n + c];
      tmp[i] = s;
    }
    for (int i = 0; i <= order; ++i) D[i * n + c] = tmp[i];
  }
}

// RHS evaluations a cell of the adaptive integrators may use: a bound on the
// work of one cell (16 per allowed step) whatever the arithmetic does, so that
// every per-cell loop ends.
PPC_DI long long rhs_budget(int max_steps) { return 16LL * max_steps + 256; }

// Adaptive BDF for one cell from t = 0 to t_end; y (n) in/out.  LU_T is the
// precision of the Newton-matrix factorisation (double: plain FP64 form).
// Returns the accepted step count (>= 0) or a negative status.
template <class M, typename LU_T>
PPC_DI int adaptive_cell(const AdaptParamsDev& p, const int reactor_type, double* __restrict__ y0,
                         const double rhoe_init, const double rhoesrc, const double* __restrict__ rYsrc,
                         const double t_bound, int* __restrict__ cnt, const LaneWS& w) {
  constexpr int n = M::NEQ;
  constexpr bool WS = uses_ws<M>();
  const double rtol = p.rtol, atol = p.atol;
  const double newton_tol = fmax(10 * 2.220446049250313e-16 / rtol, fmin(0.03, sqrt(rtol)));
  // CVODE-style options (ReactorCvode.cpp:206-224, ReactorUtils.H:119-131);
  // 0 / null keep SciPy's order 5, 4 Newton iterations, no Jacobian age, scalar atol
  const int maxord = (p.max_order > 0 && p.max_order < AMAXORD) ? p.max_order : AMAXORD;
  const int nmax = p.newton_maxiter > 0 ? p.newton_maxiter : 4;
  const double* __restrict__ atolv = p.atolv;
  auto atol_of = [&](int i) { return atolv ? atolv[i] : atol; };
  int jac_step = 0;
  double D[(AMAXORD + 3) * n];
  double J_p[WS ? 1 : n * n];
  LU_T LU_p[WS ? 1 : n * n];
  const auto J = lane_array<WS>(J_p, w, 0);
  const auto LU = lane_array<WS>(LU_p, w, n * n);
  int piv[n];
  double ypred[n], scale[n], psi[n], yv[n], d[n], f[n], dy[n];
  int nsteps = 0, nrhs = 0, njac = 0, nlu = 0, nnewton = 0, nrej = 0;
  double t = 0.0;
  f_kernel_spec<M>(t, reactor_type, y0, f, rhoe_init, rhoesrc, rYsrc);
  nrhs++;
  double h_abs;
  if (p.first_step > 0) {
    h_abs = p.first_step;
  } else {
    // initial step (Hairer, Norsett & Wanner I, II.4)
    for (int i = 0; i < n; ++i) scale[i] = atol_of(i) + fabs(y0[i]) * rtol;
    const double d0 = rms_norm<n>(y0, scale), d1 = rms_norm<n>(f, scale);
    double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    h0 = fmin(h0, t_bound);
    for (int i = 0; i < n; ++i) yv[i] = y0[i] + h0 * f[i];
    f_kernel_spec<M>(t + h0, reactor_type, yv, dy, rhoe_init, rhoesrc, rYsrc);
    nrhs++;
    for (int i = 0; i < n; ++i) dy[i] = dy[i] - f[i];
    const double d2 = rms_norm<n>(dy, scale) / h0;
    const double h1 = (d1 <= 1e-15 && d2 <= 1e-15) ? fmax(1e-6, h0 * 1e-3)
                                                   : sqrt(0.01 / fmax(d1, d2));
    h_abs = fmin(fmin(100 * h0, h1), t_bound);
  }
  for (int i = 0; i < n; ++i) {
    D[i] = y0[i];
    D[n + i] = f[i] * h_abs;
  }
  // the higher rows are written before they are read; zeroed so that no
  // path reads indeterminate values
  for (int i = 2 * n; i < (AMAXORD + 3) * n; ++i) D[i] = 0.0;
  int order = 1, n_equal = 0;
  bool lu_valid = false;
  bdf_jacobian<M>(reactor_type, y0, 1.0, J, J);
  njac++;
  int status = 0;
  const long long rhs_cap = rhs_budget(p.max_steps);
  while (t < t_bound) {
    if (nsteps >= p.max_steps) {
      status = -2;
      break;
    }
    const double min_step = 10 * fabs(nextafter(t, INFINITY) - t);
    if (h_abs < min_step) {
      change_D<n>(D, order, min_step / h_abs);
      h_abs = min_step;
      n_equal = 0;
    }
    bool current_jac = false, accepted = false;
    int n_iter = 0;
    double t_new = t, error_norm = 0.0, safety = 0.9;
    while (!accepted) {
      if (!(h_abs >= min_step)) {  // also ends a cell whose step size is NaN
        status = -3;
        break;
      }
      if (nrhs >= rhs_cap) {
        status = -2;
        break;
      }
      t_new = t + h_abs;
      if (t_new - t_bound > 0) {
        t_new = t_bound;
        change_D<n>(D, order, fabs(t_new - t) / h_abs);
        n_equal = 0;
        lu_valid = false;
      }
      const double h = t_new - t;
      h_abs = fabs(h);
      const double inv_alpha = 1.0 / ndf_alpha(order);
      for (int i = 0; i < n; ++i) {
        double sum = 0.0, ps = 0.0;
        for (int k = 0; k <= order; ++k) sum += D[k * n + i];
        for (int k = 1; k <= order; ++k) ps += D[k * n + i] * bdf_gamma(k);
        ypred[i] = sum;
        scale[i] = atol_of(i) + rtol * fabs(sum);
        psi[i] = ps * inv_alpha;
      }
      // a Jacobian older than jac_age steps is re-evaluated at the prediction
      if (p.jac_age > 0 && !current_jac && nsteps - jac_step >= p.jac_age) {
        bdf_jacobian<M>(reactor_type, ypred, 1.0, J, J);
        njac++;
        lu_valid = false;
        current_jac = true;
        jac_step = nsteps;
      }
      bool converged = false;
      const double c = h * inv_alpha;
      while (!converged) {
        if (!lu_valid) {
          for (int i0 = 0; i0 < n * n; i0 += LU_CHUNK) {  // loads of a chunk before its stores
            double v[LU_CHUNK];
#pragma unroll
            for (int u = 0; u < LU_CHUNK; ++u)
              if (i0 + u < n * n) v[u] = J[i0 + u];
#pragma unroll
            for (int u = 0; u < LU_CHUNK; ++u)
              if (i0 + u < n * n) LU[i0 + u] = (LU_T)(-c * v[u]);
          }
          for (int i = 0; i < n; ++i) LU[i * n + i] += (LU_T)1;
          lu_factor<n, LU_T>(LU, piv);
          nlu++;
          lu_valid = true;
        }
        for (int i = 0; i < n; ++i) {
          yv[i] = ypred[i];
          d[i] = 0.0;
        }
        double dy_norm_old = -1.0;
        int k;
        for (k = 0; k < nmax; ++k) {
          f_kernel_spec<M>(t_new, reactor_type, yv, f, rhoe_init, rhoesrc, rYsrc);
          nrhs++;
          nnewton++;
          bool finite = true;
          for (int i = 0; i < n; ++i) finite = finite && isfinite(f[i]);
          if (!finite) break;
          for (int i = 0; i < n; ++i) dy[i] = c * f[i] - psi[i] - d[i];
          lu_solve<n, LU_T>(LU, piv, dy);
          const double dy_norm = rms_norm<n>(dy, scale);
          const double rate = dy_norm_old < 0 ? -1.0 : dy_norm / dy_norm_old;
          if (rate >= 0 && (rate >= 1 || pow(rate, nmax - k) / (1 - rate) * dy_norm > newton_tol)) break;
          for (int i = 0; i < n; ++i) {
            yv[i] += dy[i];
            d[i] += dy[i];
          }
          if (dy_norm == 0 || (rate >= 0 && rate / (1 - rate) * dy_norm < newton_tol)) {
            converged = true;
            break;
          }
          dy_norm_old = dy_norm;
        }
        n_iter = k + 1;
        if (!converged) {
          if (current_jac) break;
          bdf_jacobian<M>(reactor_type, ypred, 1.0, J, J);
          njac++;
          lu_valid = false;
          current_jac = true;
          jac_step = nsteps;
        }
      }
      if (!converged) {
        h_abs *= 0.5;
        change_D<n>(D, order, 0.5);
        n_equal = 0;
        lu_valid = false;
        nrej++;
        continue;
      }
      safety = 0.9 * (2 * nmax + 1) / (double)(2 * nmax + n_iter);
      for (int i = 0; i < n; ++i) {
        scale[i] = atol_of(i) + rtol * fabs(yv[i]);
        dy[i] = ndf_econst(order) * d[i];
      }
      error_norm = rms_norm<n>(dy, scale);
      if (error_norm > 1) {
        const double factor = fmax(0.2, safety * pow(error_norm, -1.0 / (order + 1)));
        h_abs *= factor;
        change_D<n>(D, order, factor);
        n_equal = 0;
        nrej++;
      } else {
        accepted = true;
      }
    }
    if (status != 0) break;
    nsteps++;
    n_equal++;
    t = t_new;
    for (int i = 0; i < n; ++i) {
      D[(order + 2) * n + i] = d[i] - D[(order + 1) * n + i];
      D[(order + 1) * n + i] = d[i];
    }
    for (int k = order; k >= 0; --k)
      for (int i = 0; i < n; ++i) D[k * n + i] += D[(k + 1) * n + i];
    if (n_equal < order + 1) continue;
    double em = INFINITY, ep = INFINITY;
    if (order > 1) {
      for (int i = 0; i < n; ++i) dy[i] = ndf_econst(order - 1) * D[order * n + i];
      em = rms_norm<n>(dy, scale);
    }
    if (order < maxord) {
      for (int i = 0; i < n; ++i) dy[i] = ndf_econst(order + 1) * D[(order + 2) * n + i];
      ep = rms_norm<n>(dy, scale);
    }
    const double norms[3] = {em, error_norm, ep};
    double best = -1.0;
    int delta = 0;
    for (int i = 0; i < 3; ++i) {
      const double fct = (norms[i] == 0.0) ? INFINITY : pow(norms[i], -1.0 / (order + i));
      if (fct > best) {
        best = fct;
        delta = i - 1;
      }
    }
    order += delta;
    const double factor = fmin(10.0, safety * best);
    h_abs *= factor;
    change_D<n>(D, order, factor);
    n_equal = 0;
    lu_valid = false;
  }
  for (int i = 0; i < n; ++i) y0[i] = D[i];
  cnt[0] = nrhs;
  cnt[1] = njac;
  cnt[2] = nlu;
  cnt[3] = nnewton;
  cnt[4] = nrej;
  return status == 0 ? nsteps : status;
}

// -------------------------------------------------------------- kernels --
template <class M, int MODE>
PPC_DI int integrate_cell(const ReactParamsDev& P, double* y, double rhoe, double rhoesrc,
                          const double* src, double dt, int* cnt, const LaneWS& w) {
  if constexpr (MODE == MODE_ERK) {
    return erk_cell<M>(P, P.reactor_type, y, rhoe, rhoesrc, src, dt, cnt, w);
  } else {
    if (MODE == MODE_BDF) return bdf_parity_cell<M>(P.bdf, P.reactor_type, y, rhoe, rhoesrc, src, dt, w);
    if (MODE == MODE_RK64) return rk64_cell<M>(P.rk64, P.reactor_type, y, rhoe, rhoesrc, src, dt);
    return adaptive_cell<M, double>(P.adapt, P.reactor_type, y, rhoe, rhoesrc, src, dt, cnt, w);
  }
}

PPC_DI void store_counters(const CellStats& st, int c, int nsteps, const int* cnt) {
  st.nsteps[c] = nsteps;
  if (st.nrhs) {
    st.nrhs[c] = cnt[0];
    st.njac[c] = cnt[1];
    st.nlu[c] = cnt[2];
    st.nnewton[c] = cnt[3];
    st.nreject[c] = cnt[4];
  }
}

// 1-D (YCOrder) form: ReactorBDF.cpp:133-311 / ReactorRK64.cpp:22-164.  One
// lane per cell; with a workspace (lane_ws.h) the grid is what the workspace
// holds and lanes stride over the cells.
template <class M, int MODE>
PPC_DI void react_1d_cells(const ReactParamsDev& P, double* __restrict__ yv, const double* __restrict__ ysrc,
                           double* __restrict__ rX, const double* __restrict__ rXsrc, double dt, int ncells,
                           const CellStats& st) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const LaneWS w{P.ws, P.ws_lanes, g};
  constexpr int n = M::NEQ;
  for (int c = g; c < ncells; c += gridDim.x * blockDim.x) {
    double y[n], src[M::NS];
    for (int k = 0; k < n; ++k) y[k] = yv[(long long)c * n + k];
    for (int k = 0; k < M::NS; ++k) src[k] = ysrc[(long long)c * M::NS + k];
    const double rhoe = rX[c], rhoesrc = rXsrc[c];
    int cnt[5] = {0, 0, 0, 0, 0};
    const int r = integrate_cell<M, MODE>(P, y, rhoe, rhoesrc, src, dt, cnt, w);
    for (int k = 0; k < n; ++k) yv[(long long)c * n + k] = y[k];
    rX[c] = rhoe + dt * rhoesrc;
    store_counters(st, c, r, cnt);
  }
}

// Box (Array4) form: ReactorBDF.cpp:314-494 / ReactorRK64.cpp:166-320:
// initial T from energy, integration, rE += dt*rEsrc, final T from energy with
// the integrated T as guess, FC = per-cell step count.
template <class M, int MODE>
PPC_DI void react_box_cells(const ReactParamsDev& P, const BoxArgs& b, double dt, int ncells, const CellStats& st) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const LaneWS w{P.ws, P.ws_lanes, g};
  constexpr int NS = M::NS, n = M::NEQ;
  for (int c = g; c < ncells; c += gridDim.x * blockDim.x) {
    const int i = b.lo[0] + c % b.len[0];
    const int j = b.lo[1] + (c / b.len[0]) % b.len[1];
    const int k = b.lo[2] + c / (b.len[0] * b.len[1]);
    double y[n], src[NS], Y[NS];
    for (int s = 0; s < NS; ++s) {
      y[s] = b.rY(i, j, k, s);
      src[s] = b.rYsrc(i, j, k, s);
    }
    double rinv;
    double rho = rho_and_Y<M>(y, Y, rinv);
    const double rhoe = b.rE(i, j, k, 0), rhoesrc = b.rEsrc(i, j, k, 0);
    y[NS] = T_from_energy_rt<M>(P.reactor_type, rhoe / rho, Y, b.T(i, j, k, 0));
    int cnt[5] = {0, 0, 0, 0, 0};
    const int r = integrate_cell<M, MODE>(P, y, rhoe, rhoesrc, src, dt, cnt, w);
    rho = rho_and_Y<M>(y, Y, rinv);
    for (int s = 0; s < NS; ++s) b.rY(i, j, k, s) = y[s];
    const double rE_new = rhoe + dt * rhoesrc;
    b.rE(i, j, k, 0) = rE_new;
    b.T(i, j, k, 0) = T_from_energy_rt<M>(P.reactor_type, rE_new / rho, Y, y[NS]);
    if (b.FC.p) b.FC(i, j, k, 0) = (double)(P.adapt.fc_rhs && r >= 0 ? cnt[0] : r);
    store_counters(st, c, r, cnt);
  }
}

// The kernels at the compiler's choice of occupancy ...
template <class M, int MODE>
__global__ void __launch_bounds__(64) react_1d_kernel(ReactParamsDev P, double* __restrict__ yv,
                                                      const double* __restrict__ ysrc,
                                                      double* __restrict__ rX,
                                                      const double* __restrict__ rXsrc, double dt,
                                                      int ncells, CellStats st) {
  react_1d_cells<M, MODE>(P, yv, ysrc, rX, rXsrc, dt, ncells, st);
}

template <class M, int MODE>
__global__ void __launch_bounds__(64) react_box_kernel(ReactParamsDev P, BoxArgs b, double dt, int ncells,
                                                       CellStats st) {
  react_box_cells<M, MODE>(P, b, dt, ncells, st);
}

// ... and held to OCC waves per SIMD (2: 256 VGPRs, 4: 128 VGPRs per lane).
template <class M, int MODE, int OCC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
react_1d_kernel_held(ReactParamsDev P, double* __restrict__ yv, const double* __restrict__ ysrc,
                     double* __restrict__ rX, const double* __restrict__ rXsrc, double dt, int ncells,
                     CellStats st) {
  react_1d_cells<M, MODE>(P, yv, ysrc, rX, rXsrc, dt, ncells, st);
}

template <class M, int MODE, int OCC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
react_box_kernel_held(ReactParamsDev P, BoxArgs b, double dt, int ncells, CellStats st) {
  react_box_cells<M, MODE>(P, b, dt, ncells, st);
}

}  // namespace ppc
