// Batched adaptive BDF: persistent waves, a global work queue, and an
// integrator written as a per-lane state machine around the right-hand side.
//
// Every trip through the main loop evaluates one RHS for every lane that
// still has a cell and is not waiting for a factorisation (below), then
// advances that lane's integrator (initial step selection, modified-Newton
// iteration, error test, order/step selection, Jacobian and LU requests)
// until it needs the next RHS.  Lanes are never idle while a neighbour of the
// same wavefront is still integrating a stiffer cell: when a lane finishes its
// cell it stores the result and takes the next cell from the queue (atomic
// counter over a cost-ordered cell list).
//
// The algorithm follows SciPy's scipy/integrate/_ivp/bdf.py (BSD-3-Clause,
// Copyright (c) 2001-2002 Enthought, Inc. 2003-2024, SciPy Developers; the
// notice and license are in THIRD_PARTY_NOTICES.md at the repository root).
// The step/order/Newton logic is the same NDF family as adaptive_cell
// (reactors.h); the Newton matrix I - cJ is equilibrated with the error
// weights, S^-1 (I - cJ) S, and factored in LU_T (float: FP32 LU with FP64
// residuals, rates, T inversion and error control).
//
// A wave executes a Jacobian or an LU factorisation whenever any of its lanes
// needs one, so with one request per lane every ~10 trips nearly every trip
// would pay for a factorisation.  Lanes therefore post Jacobian / LU requests
// and the wave serves them together, at one place in the trip, once
// lu_batch lanes wait, every remaining lane waits, or a request has waited
// lu_wait trips; a waiting lane skips its trips (no RHS, no Newton
// iteration) until it is served.  The arithmetic of every cell is unchanged,
// so results do not depend on the batching.
#pragma once
#include "reactors.h"

namespace ppc {

// Cell input/output of the 1-D (YCOrder) form.
struct IO1D {
  double* y;
  const double* ysrc;
  double* rX;
  const double* rXsrc;
  template <class M>
  PPC_DI void load(int c, double* y0, double* src, double& rhoe, double& rhoesrc, int) const {
    for (int k = 0; k < M::NEQ; ++k) y0[k] = y[(long long)c * M::NEQ + k];
    for (int k = 0; k < M::NS; ++k) src[k] = ysrc[(long long)c * M::NS + k];
    rhoe = rX[c];
    rhoesrc = rXsrc[c];
  }
  template <class M>
  PPC_DI void store(int c, const double* yv, double rhoe, double rhoesrc, double dt, int nsteps, int) const {
    for (int k = 0; k < M::NEQ; ++k) y[(long long)c * M::NEQ + k] = yv[k];
    rX[c] = rhoe + dt * rhoesrc;
  }
};

// Cell input/output of the box (Array4) form, with the initial and final
// temperature inversions of ReactorBDF.cpp:376-389,461-478.
struct IOBox {
  BoxArgs b;
  PPC_DI void ijk(int c, int& i, int& j, int& k) const {
    i = b.lo[0] + c % b.len[0];
    j = b.lo[1] + (c / b.len[0]) % b.len[1];
    k = b.lo[2] + c / (b.len[0] * b.len[1]);
  }
  template <class M>
  PPC_DI void load(int c, double* y0, double* src, double& rhoe, double& rhoesrc, int reactor_type) const {
    int i, j, k;
    ijk(c, i, j, k);
    double Y[M::NS], rinv;
    for (int s = 0; s < M::NS; ++s) {
      y0[s] = b.rY(i, j, k, s);
      src[s] = b.rYsrc(i, j, k, s);
    }
    const double rho = rho_and_Y<M>(y0, Y, rinv);
    rhoe = b.rE(i, j, k, 0);
    rhoesrc = b.rEsrc(i, j, k, 0);
    y0[M::NS] = T_from_energy_rt<M>(reactor_type, rhoe / rho, Y, b.T(i, j, k, 0));
  }
  template <class M>
  PPC_DI void store(int c, const double* yv, double rhoe, double rhoesrc, double dt, int nsteps,
                    int reactor_type) const {
    int i, j, k;
    ijk(c, i, j, k);
    double Y[M::NS], rinv;
    const double rho = rho_and_Y<M>(yv, Y, rinv);
    for (int s = 0; s < M::NS; ++s) b.rY(i, j, k, s) = yv[s];
    const double rE_new = rhoe + dt * rhoesrc;
    b.rE(i, j, k, 0) = rE_new;
    b.T(i, j, k, 0) = T_from_energy_rt<M>(reactor_type, rE_new / rho, Y, yv[M::NS]);
    if (b.FC.p) b.FC(i, j, k, 0) = (double)nsteps;
  }
};

enum : int { PH_F0 = 0, PH_F1 = 1, PH_NEWTON = 2 };

// OCC: waves per SIMD the register allocation is held to.  At 2 (256 VGPRs)
// more of the per-lane state spills but a second wave hides private-memory
// latency; at 1 (256 VGPRs + 256 AGPRs) a third of the spills remain.  Which
// wins depends on the mechanism and on how often the lanes factor (drm19
// 64^3, 1-D form: 277 ms at 2, 231 ms at 1); both are compiled and the
// launch picks one (pp_chem_config.batched_waves_per_simd).
template <class M, typename LU_T, class IO, int OCC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(OCC, OCC == 1 ? 1 : 8))) batched_kernel(ReactParamsDev P, IO io, double dt, QueueArgs q,
                                                     CellStats st) {
  constexpr int n = M::NEQ, NS = M::NS;
  const AdaptParamsDev& ap = P.adapt;
  const double rtol = ap.rtol, atol = ap.atol;
  const double newton_tol = fmax(10 * 2.220446049250313e-16 / rtol, fmin(0.03, sqrt(rtol)));
  const double t_bound = dt;

  // per-lane integrator state; the matrices of the larger mechanisms live in
  // the workspace (lane_ws.h), rows [LU | Jf | Jd]
  constexpr bool WS = uses_ws<M>();
  const LaneWS w{P.ws, P.ws_lanes, (long long)blockIdx.x * blockDim.x + threadIdx.x};
  double D[(AMAXORD + 3) * n];
  LU_T LU_p[WS ? 1 : n * n];
  float Jf_p[WS ? 1 : n * n];
  double Jd_p[WS ? 1 : n * n];
  const auto LU = lane_array<WS>(LU_p, w, 0);
  const auto Jf = lane_array<WS>(Jf_p, w, ws_rows<LU_T>(n * n));  // Jacobian kept for LU refreshes (FP32)
  const auto Jd = lane_array<WS>(Jd_p, w, ws_rows<LU_T>(n * n) + ws_rows<float>(n * n));  // FP64 evaluation
  int piv[n];
  double yv[n], f[n], d[n], ypred[n], psi[n], scl[n], lus[n], src[NS];
  double rhoe = 0, rhoesrc = 0;
  double t = 0, t_new = 0, h_abs = 0, c = 0, dy_norm_old = -1, h0 = 0;
  int order = 1, n_equal = 0, k_newton = 0, phase = PH_F0;
  bool lu_valid = false, current_jac = false;
  int nsteps = 0, nrhs = 0, njac = 0, nlu = 0, nnewton = 0, nrej = 0, status = 0;
  // pending Jacobian (at yjac) and LU requests, trips waited
  bool need_jac = false, need_lu = false;
  int wait = 0;
  double yjac[n];
  const int lu_batch = ap.lu_batch > 0 ? ap.lu_batch : 64, lu_wait = ap.lu_wait > 0 ? ap.lu_wait : 32;
  // RHS budget per cell: every path either accepts a step (at most max_steps)
  // or shrinks h until the minimum-step test ends the cell; the budget bounds
  // the lane's work whatever the arithmetic does
  const long long rhs_cap = rhs_budget(ap.max_steps);

  int cell = 0, cid = 0;
  bool active = false;

  auto start_cell = [&]() {
    cid = q.order ? q.order[cell] : cell;
    io.template load<M>(cid, yv, src, rhoe, rhoesrc, P.reactor_type);
    for (int i = 0; i < n; ++i) D[i] = yv[i];
    // the higher rows are written before they are read; zeroed so that no
    // path reads indeterminate values or a previous cell's
    for (int i = n; i < (AMAXORD + 3) * n; ++i) D[i] = 0.0;
    t = 0.0;
    phase = PH_F0;
    order = 1;
    n_equal = 0;
    lu_valid = false;
    need_jac = need_lu = false;
    wait = 0;
    nsteps = nrhs = njac = nlu = nnewton = nrej = status = 0;
  };

  auto finish_cell = [&]() {
    io.template store<M>(cid, D, rhoe, rhoesrc, dt, status == 0 ? nsteps : status, P.reactor_type);
    st.nsteps[cid] = status == 0 ? nsteps : status;
    if (st.nrhs) {
      st.nrhs[cid] = nrhs;
      st.njac[cid] = njac;
      st.nlu[cid] = nlu;
      st.nnewton[cid] = nnewton;
      st.nreject[cid] = nrej;
    }
    cell = atomicAdd(q.counter, 1);
    active = cell < q.ncells;
    if (active) start_cell();
  };

  cell = atomicAdd(q.counter, 1);
  active = cell < q.ncells;
  if (active) start_cell();

  // Factor S^-1 (I - c J) S in LU_T with S = current error weights.
  auto factor = [&]() {
    for (int i = 0; i < n; ++i) lus[i] = scl[i];
    for (int i = 0; i < n; ++i) {
      const double si = 1.0 / lus[i];
      for (int j0 = 0; j0 < n; j0 += LU_CHUNK) {  // loads of a chunk before its stores
        float jf[LU_CHUNK];
#pragma unroll
        for (int u = 0; u < LU_CHUNK; ++u)
          if (j0 + u < n) jf[u] = Jf[i * n + j0 + u];
#pragma unroll
        for (int u = 0; u < LU_CHUNK; ++u)
          if (j0 + u < n) {
            const int j = j0 + u;
            const double m = (i == j ? 1.0 : 0.0) - c * (double)jf[u];
            LU[i * n + j] = (LU_T)(si * m * lus[j]);
          }
      }
    }
    lu_factor<n, LU_T>(LU, piv);
    nlu++;
    lu_valid = true;
  };

  auto jacobian = [&](const double* yj) {
    bdf_jacobian<M>(P.reactor_type, yj, 1.0, Jd, Jd);
    for (int i0 = 0; i0 < n * n; i0 += LU_CHUNK) {
      double v[LU_CHUNK];
#pragma unroll
      for (int u = 0; u < LU_CHUNK; ++u)
        if (i0 + u < n * n) v[u] = Jd[i0 + u];
#pragma unroll
      for (int u = 0; u < LU_CHUNK; ++u)
        if (i0 + u < n * n) Jf[i0 + u] = (float)v[u];
    }
    njac++;
  };

  auto request_jacobian = [&](const double* yj) {
    for (int i = 0; i < n; ++i) yjac[i] = yj[i];
    need_jac = true;
  };

  // top of one step attempt: predictor, weights, LU if needed
  auto begin_attempt = [&]() -> bool {
    const double min_step = 10 * fabs(nextafter(t, INFINITY) - t);
    if (!(h_abs >= min_step)) {  // also ends a cell whose step size is NaN
      status = -3;
      return false;
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
      for (int kk = 0; kk <= order; ++kk) sum += D[kk * n + i];
      for (int kk = 1; kk <= order; ++kk) ps += D[kk * n + i] * bdf_gamma(kk);
      ypred[i] = sum;
      scl[i] = atol + rtol * fabs(sum);
      psi[i] = ps * inv_alpha;
    }
    c = h * inv_alpha;
    if (!lu_valid) need_lu = true;
    for (int i = 0; i < n; ++i) {
      yv[i] = ypred[i];
      d[i] = 0.0;
    }
    dy_norm_old = -1.0;
    k_newton = 0;
    phase = PH_NEWTON;
    return true;
  };

  auto begin_step = [&]() -> bool {
    if (nsteps >= ap.max_steps) {
      status = -2;
      return false;
    }
    const double min_step = 10 * fabs(nextafter(t, INFINITY) - t);
    if (h_abs < min_step) {
      change_D<n>(D, order, min_step / h_abs);
      h_abs = min_step;
      n_equal = 0;
    }
    current_jac = false;
    return begin_attempt();
  };

  while (active) {
    // ---- Jacobian / LU requests, served for the waiting lanes together
    const unsigned long long waiting = __ballot(need_lu);
    if (waiting) {
      const unsigned long long all = __ballot(true);
      const bool stale = __ballot(need_lu && wait >= lu_wait) != 0;
      if (__popcll(waiting) >= lu_batch || waiting == all || stale) {
        if (need_jac) {
          jacobian(yjac);
          need_jac = false;
        }
        if (need_lu) {
          factor();
          need_lu = false;
          wait = 0;
        }
      }
    }
    if (need_lu) {  // not served yet: no RHS, no Newton iteration this trip
      ++wait;
      continue;
    }
    if (nrhs >= rhs_cap) {  // out of budget: the cell fails with the step-limit status
      status = -2;
      finish_cell();
      continue;
    }
    // ---- the one RHS evaluation of this trip
    const double t_eval = (phase == PH_F0) ? 0.0 : (phase == PH_F1 ? h0 : t_new);
    f_kernel_spec<M>(t_eval, P.reactor_type, yv, f, rhoe, rhoesrc, src);
    nrhs++;
    bool ok = true;
    if (phase == PH_F0) {
      // f0 = f(y0): initial step size (Hairer, Norsett & Wanner I, II.4)
      for (int i = 0; i < n; ++i) {
        D[n + i] = f[i];  // f0 kept in D[1] until h is known
        scl[i] = atol + fabs(D[i]) * rtol;
      }
      if (ap.first_step > 0) {
        h_abs = ap.first_step;
        for (int i = 0; i < n; ++i) D[n + i] *= h_abs;
        for (int i = 0; i < n; ++i) yv[i] = D[i];
        request_jacobian(yv);
        ok = begin_step();
      } else {
        const double d0 = rms_norm<n>(D, scl), d1 = rms_norm<n>(f, scl);
        h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
        h0 = fmin(h0, t_bound);
        for (int i = 0; i < n; ++i) yv[i] = D[i] + h0 * f[i];
        phase = PH_F1;
      }
    } else if (phase == PH_F1) {
      for (int i = 0; i < n; ++i) d[i] = f[i] - D[n + i];
      const double d1 = rms_norm<n>(D + n, scl);
      const double d2 = rms_norm<n>(d, scl) / h0;
      const double h1 = (d1 <= 1e-15 && d2 <= 1e-15) ? fmax(1e-6, h0 * 1e-3) : sqrt(0.01 / fmax(d1, d2));
      h_abs = fmin(fmin(100 * h0, h1), t_bound);
      for (int i = 0; i < n; ++i) D[n + i] *= h_abs;
      for (int i = 0; i < n; ++i) yv[i] = D[i];
      request_jacobian(yv);
      ok = begin_step();
    } else {
      // one modified-Newton iteration with f = f(t_new, yv)
      nnewton++;
      bool finite = true;
      for (int i = 0; i < n; ++i) finite = finite && isfinite(f[i]);
      bool failed = !finite, converged = false;
      int n_iter = k_newton + 1;
      if (finite) {
        double dy[n];
        for (int i = 0; i < n; ++i) dy[i] = (c * f[i] - psi[i] - d[i]) / lus[i];
        lu_solve<n, LU_T>(LU, piv, dy);
        for (int i = 0; i < n; ++i) dy[i] *= lus[i];
        const double dy_norm = rms_norm<n>(dy, scl);
        const double rate = dy_norm_old < 0 ? -1.0 : dy_norm / dy_norm_old;
        if (rate >= 0 && (rate >= 1 || pow(rate, 4 - k_newton) / (1 - rate) * dy_norm > newton_tol)) {
          failed = true;
        } else {
          for (int i = 0; i < n; ++i) {
            yv[i] += dy[i];
            d[i] += dy[i];
          }
          if (dy_norm == 0 || (rate >= 0 && rate / (1 - rate) * dy_norm < newton_tol)) {
            converged = true;
          } else {
            dy_norm_old = dy_norm;
            k_newton++;
            if (k_newton == 4) failed = true;
          }
        }
      }
      if (failed) {
        if (!current_jac) {
          // refresh the Jacobian at the predictor and retry the same step
          request_jacobian(ypred);
          current_jac = true;
          need_lu = true;
          for (int i = 0; i < n; ++i) {
            yv[i] = ypred[i];
            d[i] = 0.0;
          }
          dy_norm_old = -1.0;
          k_newton = 0;
        } else {
          h_abs *= 0.5;
          change_D<n>(D, order, 0.5);
          n_equal = 0;
          lu_valid = false;
          nrej++;
          ok = begin_attempt();
        }
      } else if (converged) {
        const double safety = 0.9 * 9.0 / (double)(8 + n_iter);
        double err[n];
        for (int i = 0; i < n; ++i) {
          scl[i] = atol + rtol * fabs(yv[i]);
          err[i] = ndf_econst(order) * d[i];
        }
        const double error_norm = rms_norm<n>(err, scl);
        if (error_norm > 1) {
          const double fac = fmax(0.2, safety * pow(error_norm, -1.0 / (order + 1)));
          h_abs *= fac;
          change_D<n>(D, order, fac);
          n_equal = 0;
          nrej++;
          ok = begin_attempt();
        } else {
          // accept
          nsteps++;
          n_equal++;
          t = t_new;
          for (int i = 0; i < n; ++i) {
            D[(order + 2) * n + i] = d[i] - D[(order + 1) * n + i];
            D[(order + 1) * n + i] = d[i];
          }
          for (int kk = order; kk >= 0; --kk)
            for (int i = 0; i < n; ++i) D[kk * n + i] += D[(kk + 1) * n + i];
          if (n_equal >= order + 1) {
            double em = INFINITY, ep = INFINITY;
            if (order > 1) {
              for (int i = 0; i < n; ++i) err[i] = ndf_econst(order - 1) * D[order * n + i];
              em = rms_norm<n>(err, scl);
            }
            if (order < AMAXORD) {
              for (int i = 0; i < n; ++i) err[i] = ndf_econst(order + 1) * D[(order + 2) * n + i];
              ep = rms_norm<n>(err, scl);
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
            const double fac = fmin(10.0, safety * best);
            h_abs *= fac;
            change_D<n>(D, order, fac);
            n_equal = 0;
            lu_valid = false;
          }
          if (t < t_bound) {
            ok = begin_step();
          } else {
            finish_cell();
            continue;
          }
        }
      }
    }
    if (!ok) finish_cell();  // failed cell: store with the negative status
  }
}

}  // namespace ppc
