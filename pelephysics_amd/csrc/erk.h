// ReactorArkode's explicit Runge-Kutta integration (ERKStep with the table of
// ode.rk_method, ReactorArkode.cpp:26-83, and the controller of
// ode.rk_controller, :85-110), one lane per cell:
//
//  * stages K_i = f(t + c_i h, y + h sum_{j<i} a_ij K_j) in the lane's rows of
//    the device workspace (s x NEQ, up to 13 stages); y_new = y + h sum b_i K_i,
//    error estimate h sum (b_i - d_i) K_i;
//  * error test: the WRMS norm with weights 1 / (rtol |y| + atol_k) (atol_k =
//    atol x typical value when set) at most 1;
//  * step factor: safety 0.96 x the controller's factor from the scaled error
//    estimates e = 1.5 x norm of this and the two accepted steps before
//    (PID: e1^-0.58/k e2^0.21/k e3^-0.1/k; PI: e1^-0.8/k e2^0.31/k; I:
//    e1^-1/k; explicit Gustafsson: e1^-0.367/k (e1/e2)^-0.268/k; k the
//    embedding's order + 1), at most 20 (10000 after the first step, 1 after
//    a failure), at least 0.1 and at most 0.3 after 2 failures in a row, kept
//    at 1 inside [1, 1.5] (ARKODE's published adaptivity defaults);
//  * first same as last: a table whose last stage is the new solution passes
//    it on as the next step's first stage; the last step is clipped to end
//    at dt_react.
// Counters: cnt[0] RHS evaluations (FC and the return value, as
// ReactorArkode.cpp:239-260 reports nfe), cnt[4] rejected steps.
#pragma once
#include "chem_cell.h"
#include "lane_ws.h"
#include "reactors_fwd.h"

namespace ppc {

PPC_DI double erk_factor(int kind, double k, double e1, double e2, double e3, double e2g) {
  switch (kind) {
    case 1: return pow(e1, -0.8 / k) * pow(e2, 0.31 / k);
    case 2: return pow(e1, -1.0 / k);
    case 3: return e2g > 0.0 ? pow(e1, -0.367 / k) * pow(e1 / e2g, -0.268 / k) : pow(e1, -1.0 / k);
    default: return pow(e1, -0.58 / k) * pow(e2, 0.21 / k) * pow(e3, -0.1 / k);
  }
}

template <int n>
PPC_DI double erk_wrms(const double* __restrict__ v, const double* __restrict__ wt) {
  double acc = 0.0;
  for (int i = 0; i < n; ++i) acc += (v[i] * wt[i]) * (v[i] * wt[i]);
  return sqrt(acc / n);
}

// Returns the accepted steps, or -2 (step limit) / -3 (step too small or not
// a number); y (NEQ) is advanced in place.
template <class M>
PPC_DI int erk_cell(const ReactParamsDev& P, const int reactor_type, double* __restrict__ y, const double rhoe,
                    const double rhoesrc, const double* __restrict__ src, const double t_end, int* cnt,
                    const LaneWS& w) {
  constexpr int n = M::NEQ;
  const ErkParamsDev& E = P.erk;
  const AdaptParamsDev& a = P.adapt;
  const int S = E.s;
  const double* __restrict__ cc = E.tab;
  const double* __restrict__ A = cc + S;
  const double* __restrict__ b = A + S * S;
  const double* __restrict__ d = b + S;
  const LaneArr<double> K = w.template at<double>(0);  // K[i * n + m]: stage i, component m
  auto rhs = [&](double t, const double* yy, long long row) {
    double f[n];
    f_kernel_spec<M>(t, reactor_type, yy, f, rhoe, rhoesrc, src);
    for (int m = 0; m < n; ++m) K[row + m] = f[m];
  };
  auto atol_of = [&](int i) { return a.atolv ? a.atolv[i] : a.atol; };
  double yt[n], wt[n], err[n];
  const double k = E.p + 1.0;
  int nst = 0, nfe = 0, nrej = 0, nef = 0, status = 0;
  bool first = true, after_fail = false;
  double t = 0.0, e2 = 1.0, e3 = 1.0, e2g = 0.0;
  rhs(t, y, 0);
  nfe++;
  bool k0_valid = true;
  // initial step (Hairer, Norsett & Wanner I, II.4) with the embedding's order
  double h;
  {
    for (int i = 0; i < n; ++i) wt[i] = 1.0 / (atol_of(i) + fabs(y[i]) * a.rtol);
    for (int i = 0; i < n; ++i) err[i] = K[i];
    const double d0 = erk_wrms<n>(y, wt), d1 = erk_wrms<n>(err, wt);
    double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    h0 = fmin(h0, t_end);
    for (int i = 0; i < n; ++i) yt[i] = y[i] + h0 * err[i];
    double f1[n];
    f_kernel_spec<M>(t + h0, reactor_type, yt, f1, rhoe, rhoesrc, src);
    nfe++;
    for (int i = 0; i < n; ++i) f1[i] = f1[i] - err[i];
    const double d2 = erk_wrms<n>(f1, wt) / h0;
    const double h1 = (d1 <= 1e-15 && d2 <= 1e-15) ? fmax(1e-6, h0 * 1e-3) : pow(0.01 / fmax(d1, d2), 1.0 / k);
    h = fmin(fmin(100 * h0, h1), t_end);
  }
  while (t < t_end) {
    if (a.max_steps > 0 && nst + nrej >= a.max_steps) {
      status = -2;
      break;
    }
    const double min_step = 10 * fabs(nextafter(t, INFINITY) - t);
    if (!(h >= min_step)) {
      status = -3;
      break;
    }
    const double hh = fmin(h, t_end - t);
    if (!k0_valid) {
      rhs(t, y, 0);
      nfe++;
    }
    for (int i = 1; i < S; ++i) {
      for (int m = 0; m < n; ++m) {
        double acc = 0.0;
        for (int j = 0; j < i; ++j) acc += A[i * S + j] * K[(long long)j * n + m];
        yt[m] = y[m] + hh * acc;
      }
      rhs(t + cc[i] * hh, yt, (long long)i * n);
      nfe++;
    }
    for (int m = 0; m < n; ++m) {
      double sb = 0.0, se = 0.0;
      for (int j = 0; j < S; ++j) {
        const double kj = K[(long long)j * n + m];
        sb += b[j] * kj;
        se += (b[j] - d[j]) * kj;
      }
      yt[m] = y[m] + hh * sb;
      err[m] = hh * se;
    }
    for (int i = 0; i < n; ++i) wt[i] = 1.0 / (atol_of(i) + fabs(y[i]) * a.rtol);
    const double dsm = erk_wrms<n>(err, wt);
    const double e1 = fmax(1.5 * dsm, 1e-10);
    double eta = 0.96 * erk_factor(E.controller, k, e1, e2, e3, e2g);
    if (dsm <= 1.0) {
      for (int m = 0; m < n; ++m) y[m] = yt[m];
      t += hh;
      nst++;
      if (E.fsal) {
        for (int m = 0; m < n; ++m) K[m] = K[(long long)(S - 1) * n + m];
      }
      k0_valid = E.fsal != 0;
      eta = fmin(eta, first ? 10000.0 : 20.0);
      if (after_fail) eta = fmin(eta, 1.0);
      if (eta >= 1.0 && eta <= 1.5) eta = 1.0;
      e3 = e2;
      e2 = e1;
      e2g = e1;
      first = false;
      nef = 0;
      after_fail = false;
      h = hh * eta;
    } else {
      nrej++;
      nef++;
      after_fail = true;
      eta = fmin(fmax(eta, 0.1), 1.0);
      if (nef >= 2) eta = fmin(eta, 0.3);
      k0_valid = true;  // K_0 = f(t, y) with y unchanged
      h = hh * eta;
    }
  }
  cnt[0] = nfe;
  cnt[4] = nrej;
  return status < 0 ? status : nst;
}

}  // namespace ppc
