// Discrete adjoint of the fixed-step BDF integration (ReactorBDF map with one
// linearised Newton iteration per substep and the linear systems solved
// exactly), one 64-lane wave per cell.
//
// Forward: nmacro calls of dt each (ReactorBDF.cpp:187-285 semantics: history
// reset at every call, TRPZ ramp, nsubsteps substeps); every substep state is
// checkpointed to global memory.  Substep map, with u = soln_n = soln:
//   r  = -((t0+t1) u + t2 nm1 + t3 nm2)/dt + fs f(u),  fs = f0 + [TRPZ] f1
//   A  = t0/dt I - f0 J(u)           (J: the BDF Jacobian of bdf_jacobian)
//   u' = u + A^-1 r,  nm1' = u,  nm2' = nm1
// Backward (recomputation from the checkpoints), for each output o with the
// adjoint lam' of u':
//   mu   = A^-T lam'
//   lam(u)   += lam' - (t0+t1)/dt mu + fs Jtrue^T mu + f0 grad_u(mu^T J(u) D)
//   lam(nm1) += -t2/dt mu,  lam(nm2) += -t3/dt mu
//   g_j      += fs mu^T df/dlnA_j + f0 mu^T (dJ/dlnA_j) D,   D = A^-1 r
// where Jtrue = df/du includes the temperature inversion of the RHS
// (dT/d(rhoY_l) = -e_l/(rho c), df/dT_integrated = 0), df/dlnA_j is reaction
// j's term of the production rates (rate constants scale with A_j; falloff
// A_inf and A_0 together), dJ/dlnA_j is reaction j's own contribution to J (J
// is linear in the reaction contributions), and grad_u(mu^T J D) is formed by
// directional derivatives of the table-driven Jacobian evaluated in dual
// numbers (one pass per component of u).
#pragma once
#include "dual.h"
#include "wave_bdf.h"

namespace ppc {

template <class T>
struct AdjChem {
  using W = WaveChem<T>;
  using TM = TabMech<T>;
  static constexpr int NS = T::NS, NR = T::NR, n = T::NEQ;
  struct Shared {
    double nasa[NS][2][7];
    double c[NS], a[NS], b[NS], q[NR];
    double g[NS], h[NS], wd[NS];
    double J[n * n];         // Jacobian core (SI, column-major) at the integrated T, reactor's consP
    double Jx[n * n];        // exact (constant-volume) core at the energy-inverted T: Jtrue
    double wdx[NS];
    double Jrow[n][n];       // A rows, then the transpose staging
    Dual cd[NS], gd[NS], hd[NS];
    double dJ[n * n], dwd[NS];
    double mu[3][n], dl[n];  // adjoint solutions and the Newton increment, broadcast
    double wa[3][NS], wb[NS], wc[3][NS], wf[3][NS];
  };

  // Jacobian core (aJacobian semantics, chem_table.h) of the cell at (sc, Tv)
  // into J / wd; with trow the temperature row too.
  template <int CONSP>
  PPC_DI static void core(const typename W::Lane& L, Shared& sh, int lane, double sc, double Tv, double* J,
                          double* wd, bool trow) {
    const double invT = 1.0 / Tv, invT2 = invT * invT, logT = log(Tv);
    const double refC = 101325.0 / 8.31446 * invT, refCinv = 1.0 / refC;
    double cf[7];
    W::coeffs(L, sh, Tv, cf);
    const double hRT = TM::h_poly(cf, Tv, invT);
    const double cpR = TM::cp_poly(cf, Tv);
    const double dcR = TM::dcp_poly(cf, Tv);
#pragma unroll 1
    for (int i = lane; i < n * n; i += 64) J[i] = 0.0;
    if (L.sp) {
      sh.c[lane] = sc;
      sh.g[lane] = TM::g_poly(cf, Tv, invT, logT);
      sh.h[lane] = hRT;
      wd[lane] = 0.0;
    }
    const double mixture = wave::sum(L.sp ? sc : 0.0);
    __syncthreads();
#pragma unroll 1
    for (int r0 = 0; r0 < NR; r0 += 64) {
      const int j = r0 + lane;
      if (j < NR)
        TM::template jac_reaction<CONSP, double>(
            j, sh.c, sh.g, sh.h, mixture, Tv, invT, invT2, logT, refC, refCinv,
            [&](int idx, double v) { atomicAdd(&J[idx], v); }, [&](int m, double v) { atomicAdd(&wd[m], v); });
    }
    __syncthreads();
    if (trow) {
      const double cR = (CONSP == 1) ? cpR : cpR - 1.0;
      const double eh = (CONSP == 1) ? hRT : hRT - 1.0;
      const double wdk = L.sp ? wd[lane] : 0.0;
      const double cmix = wave::sum(L.sp ? cR * sc : 0.0);
      const double dcmixdT = wave::sum(L.sp ? dcR * sc : 0.0);
      const double ehmix = wave::sum(L.sp ? eh * wdk : 0.0);
      const double dehmixdT = wave::sum(L.sp ? invT * (cR - eh) * wdk + eh * J[NS * n + lane] : 0.0);
      const double cmixinv = 1.0 / cmix;
      const double tmp1 = ehmix * cmixinv, tmp3 = cmixinv * Tv, tmp2 = tmp1 * tmp3;
      if (L.sp) {
        double d = 0.0;
#pragma unroll 1
        for (int mm = 0; mm < NS; ++mm) d += ((CONSP == 1) ? sh.h[mm] : sh.h[mm] - 1.0) * J[lane * n + mm];
        J[lane * n + NS] = tmp2 * cR - tmp3 * d;
      }
      if (lane == NS) J[n * n - 1] = -tmp1 + tmp2 * dcmixdT - tmp3 * dehmixdT;
      __syncthreads();
    }
  }

  // row `lane` of the BDF Jacobian (bdf_jacobian conversion of a core matrix)
  PPC_DI static double ref_entry(const typename W::Lane& L, const double* J, int i, int jj) {
    if (jj < NS) {
      const double Jv = (i < NS) ? J[jj * n + i] : J[jj * n + NS] * 1.0e6;
      return Jv * (L.sp ? L.mw : 1.0) * T::imw_(jj);
    }
    const double Jv = (i < NS) ? J[NS * n + i] * 1.0e-6 : J[n * n - 1];
    return Jv * (L.sp ? L.mw : 1.0);
  }

  // Directional derivative of the BDF Jacobian along u_l (l < NS: rhoY_l, l = NS:
  // T), contracted: returns sum_i mu_o,i sum_j dJ_ij/du_l D_j for o = 0..2.
  template <int CONSP>
  PPC_DI static void dual_contract(const typename W::Lane& L, Shared& sh, int lane, double sc, double Tv, int l,
                                   double* H) {
    const bool dT = (l == NS);
    const Dual TvD(Tv, dT ? 1.0 : 0.0);
    const Dual invT = 1.0 / TvD, invT2 = invT * invT, logT = log(TvD);
    const Dual refC = (101325.0 / 8.31446) * invT, refCinv = 1.0 / refC;
    double cf[7];
    W::coeffs(L, sh, Tv, cf);
    const Dual scD(sc, (!dT && lane == l) ? L.imw * 1.0e6 : 0.0);
    const Dual hRT = TM::h_poly(cf, TvD, invT);
    const Dual cpR = TM::cp_poly(cf, TvD);
    const Dual dcR = TM::dcp_poly(cf, TvD);
#pragma unroll 1
    for (int i = lane; i < n * n; i += 64) sh.dJ[i] = 0.0;
    if (L.sp) {
      sh.cd[lane] = scD;
      sh.gd[lane] = TM::g_poly(cf, TvD, invT, logT);
      sh.hd[lane] = hRT;
      sh.dwd[lane] = 0.0;
    }
    const double dmix = dT ? 0.0 : wave::rdlane(L.imw, l) * 1.0e6;
    const Dual mixture(wave::sum(L.sp ? sc : 0.0), dmix);
    __syncthreads();
#pragma unroll 1
    for (int r0 = 0; r0 < NR; r0 += 64) {
      const int j = r0 + lane;
      if (j < NR)
        TM::template jac_reaction<CONSP, Dual>(
            j, sh.cd, sh.gd, sh.hd, mixture, TvD, invT, invT2, logT, refC, refCinv,
            [&](int idx, const Dual& v) { atomicAdd(&sh.dJ[idx], v.d); },
            [&](int m, const Dual& v) { atomicAdd(&sh.dwd[m], v.d); });
    }
    __syncthreads();
    // temperature row in dual numbers: values from the primal core sh.J / sh.wd
    {
      const Dual cR = (CONSP == 1) ? cpR : cpR - 1.0;
      const Dual eh = (CONSP == 1) ? hRT : hRT - 1.0;
      const Dual wdk = L.sp ? Dual(sh.wd[lane], sh.dwd[lane]) : Dual(0.0);
      const Dual sck = L.sp ? scD : Dual(0.0);
      const Dual Jtk = L.sp ? Dual(sh.J[NS * n + lane], sh.dJ[NS * n + lane]) : Dual(0.0);
      auto dsum = [](const Dual& x) { return Dual(wave::sum(x.v), wave::sum(x.d)); };
      const Dual cmix = dsum(L.sp ? cR * sck : Dual(0.0));
      const Dual dcmixdT = dsum(L.sp ? dcR * sck : Dual(0.0));
      const Dual ehmix = dsum(L.sp ? eh * wdk : Dual(0.0));
      const Dual dehmixdT = dsum(L.sp ? invT * (cR - eh) * wdk + eh * Jtk : Dual(0.0));
      const Dual cmixinv = 1.0 / cmix;
      const Dual tmp1 = ehmix * cmixinv, tmp3 = cmixinv * TvD, tmp2 = tmp1 * tmp3;
      double drow = 0.0;
      if (L.sp) {
        Dual d = 0.0;
#pragma unroll 1
        for (int mm = 0; mm < NS; ++mm) {
          const Dual ehm = (CONSP == 1) ? sh.hd[mm] : sh.hd[mm] - 1.0;
          d += ehm * Dual(sh.J[lane * n + mm], sh.dJ[lane * n + mm]);
        }
        drow = (tmp2 * cR - tmp3 * d).d;
      }
      const double dTT = (-tmp1 + tmp2 * dcmixdT - tmp3 * dehmixdT).d;
      __syncthreads();
      if (L.sp) sh.dJ[lane * n + NS] = drow;
      if (lane == NS) sh.dJ[n * n - 1] = dTT;
      __syncthreads();
    }
    double s = 0.0;
    if (lane < n) {
#pragma unroll 1
      for (int jj = 0; jj < n; ++jj) s += ref_entry(L, sh.dJ, lane, jj) * sh.dl[jj];
    }
#pragma unroll
    for (int o = 0; o < 3; ++o) H[o] = wave::sum(lane < n ? sh.mu[o][lane] * s : 0.0);
    __syncthreads();
  }

  // Per reaction lane: C_j,o = mu_o^T (dJ/dlnA_j) D from reaction j's own
  // contributions to the core (weights wa, wb, wc prepared in LDS).
  template <int CONSP>
  PPC_DI static void reaction_contract(Shared& sh, int lane, double Tv, double (&C)[2][3]) {
    const double invT = 1.0 / Tv, invT2 = invT * invT, logT = log(Tv);
    const double refC = 101325.0 / 8.31446 * invT, refCinv = 1.0 / refC;
    double mixture = 0.0;
#pragma unroll 1
    for (int k = 0; k < NS; ++k) mixture += sh.c[k];
    const double DT = sh.dl[NS];
#pragma unroll
    for (int rr = 0; rr < 2; ++rr)
#pragma unroll
      for (int o = 0; o < 3; ++o) C[rr][o] = 0.0;
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
      const int j = rr * 64 + lane;
      if (rr * 64 < NR && j < NR) {
        double c0 = 0.0, c1 = 0.0, c2 = 0.0;
        TM::template jac_reaction<CONSP, double>(
            j, sh.c, sh.g, sh.h, mixture, Tv, invT, invT2, logT, refC, refCinv,
            [&](int idx, double v) {
              const int k = idx / n, m = idx % n;
              const double w = (k < NS) ? sh.wb[k] : DT * 1.0e-6;
              c0 += v * w * sh.wa[0][m];
              c1 += v * w * sh.wa[1][m];
              c2 += v * w * sh.wa[2][m];
            },
            [&](int m, double v) {
              c0 += v * sh.wc[0][m];
              c1 += v * sh.wc[1][m];
              c2 += v * sh.wc[2][m];
            });
        C[rr][0] = c0;
        C[rr][1] = c1;
        C[rr][2] = c2;
      }
    }
  }
};

// The linear map of one substep: A = t0/dt I - f0 J(u) rows into LU and its LU
// factors (lu_* outputs).  J rows from the core sh.J.
template <class T>
PPC_DI void adj_factor(const typename WaveChem<T>::Lane& L, typename AdjChem<T>::Shared& sh, int lane,
                       bool transpose, double t0dt, double f0, double (&LU)[T::NEQ], int& kstep, int& pivlane,
                       double& dinv) {
  constexpr int n = T::NEQ;
  using A = AdjChem<T>;
  if (lane < n) {
#pragma unroll 1
    for (int jj = 0; jj < n; ++jj)
      sh.Jrow[lane][jj] = (lane == jj ? t0dt : 0.0) - f0 * A::ref_entry(L, sh.J, lane, jj);
  }
  __syncthreads();
#pragma unroll
  for (int jj = 0; jj < n; ++jj) LU[jj] = lane < n ? (transpose ? sh.Jrow[jj][lane] : sh.Jrow[lane][jj]) : 0.0;
  __syncthreads();
  wave_lu_factor<n, double>(LU, lane, kstep, pivlane, dinv);
}

// Forward with checkpoints, then the reverse sweep, for one cell.
template <class T>
PPC_DI void wave_adjoint_cell(const ReactParamsDev& P, const AdjointArgs& a, const int cell,
                              typename AdjChem<T>::Shared& sh, const typename WaveChem<T>::Lane& L,
                              const int lane) {
  using W = WaveChem<T>;
  using AC = AdjChem<T>;
  constexpr int NS = T::NS, NR = T::NR, n = T::NEQ;
  static_assert(NR <= 128, "two reactions per lane");
  const bool act = lane < n;
  const int rt = P.reactor_type;
  const int nsub = P.bdf.nsubsteps, scheme = P.bdf.scheme;
  const int first = (scheme >= 2) ? 1 : scheme, change = scheme - 2;
  const double dts = a.dt / (double)nsub;
  const int nstep = a.nmacro * nsub;
  {
    const double rhoe = a.rX[cell];
    double* ck = a.ckpt + (long long)cell * nstep * n;
    double u = act ? a.y[(long long)cell * n + lane] : 0.0;
    double LU[n];
    int kstep, pivlane;
    double dinv;
    // ---------------- forward, with checkpoints
    for (int c = 0, s = 0; c < a.nmacro; ++c) {
      double nm1 = u, nm2 = u;
      for (int st = 0; st < nsub; ++st, ++s) {
        const int ts = (st > change) ? scheme : first;
        const double t0 = tcoeff(ts, 0), t1 = tcoeff(ts, 1), t2 = tcoeff(ts, 2), t3 = tcoeff(ts, 3);
        const double f0 = fcoeff(ts, 0), fs = f0 + (ts == 1 ? fcoeff(ts, 1) : 0.0);
        if (act) ck[(long long)s * n + lane] = u;
        const double f = W::rhs(L, sh, lane, st * dts, rt, u, 0.0, rhoe, 0.0);
        const double r = -((t0 + t1) * u + t2 * nm1 + t3 * nm2) / dts + fs * f;
        const double sc = L.sp ? u * L.imw * 1.0e6 : 0.0;
        const double Tv = wave::rdlane(u, NS);
        if (rt == H_REACTOR)
          AC::template core<1>(L, sh, lane, sc, Tv, sh.J, sh.wd, true);
        else
          AC::template core<0>(L, sh, lane, sc, Tv, sh.J, sh.wd, true);
        adj_factor<T>(L, sh, lane, false, t0 / dts, f0, LU, kstep, pivlane, dinv);
        const double D = wave_lu_solve<n, double>(LU, lane, kstep, pivlane, dinv, act ? r : 0.0);
        nm2 = nm1;
        nm1 = u;
        u += D;
      }
    }
    if (act) a.y[(long long)cell * n + lane] = u;
    // outputs and their gradients with respect to the final state
    const double ssum = wave::sum(L.sp ? u : 0.0);
    const double ua = wave::rdlane(u, a.spec_a), ub = wave::rdlane(u, a.spec_b);
    const double Tfin = wave::rdlane(u, NS);
    if (lane == 0) {
      a.out[(long long)cell * 3 + 0] = Tfin;
      a.out[(long long)cell * 3 + 1] = ua / ssum;
      a.out[(long long)cell * 3 + 2] = ub / ssum;
    }
    // lam[o][slot]: slot 0 = state of the current substep's output position
    double lam[3][3];
    lam[0][0] = (lane == NS) ? 1.0 : 0.0;
    lam[1][0] = L.sp ? (-ua / (ssum * ssum) + (lane == a.spec_a ? 1.0 / ssum : 0.0)) : 0.0;
    lam[2][0] = L.sp ? (-ub / (ssum * ssum) + (lane == a.spec_b ? 1.0 / ssum : 0.0)) : 0.0;
#pragma unroll
    for (int o = 0; o < 3; ++o) lam[o][1] = lam[o][2] = 0.0;
    double g[2][3] = {{0, 0, 0}, {0, 0, 0}};
    // ---------------- backward
    for (int c = a.nmacro - 1; c >= 0; --c) {
      for (int st = nsub - 1; st >= 0; --st) {
        const int s = c * nsub + st;
        const int ts = (st > change) ? scheme : first;
        const double t0 = tcoeff(ts, 0), t1 = tcoeff(ts, 1), t2 = tcoeff(ts, 2), t3 = tcoeff(ts, 3);
        const double f0 = fcoeff(ts, 0), fs = f0 + (ts == 1 ? fcoeff(ts, 1) : 0.0);
        // adjoint of this substep's output u_{st+1}: slot 0; shift the pending slots
        double lo[3];
#pragma unroll
        for (int o = 0; o < 3; ++o) {
          lo[o] = lam[o][0];
          lam[o][0] = lam[o][1];
          lam[o][1] = lam[o][2];
          lam[o][2] = 0.0;
        }
        const double us = act ? ck[(long long)s * n + lane] : 0.0;
        const double nm1 = act ? ck[(long long)(c * nsub + (st >= 1 ? st - 1 : 0)) * n + lane] : 0.0;
        const double nm2 = act ? ck[(long long)(c * nsub + (st >= 2 ? st - 2 : 0)) * n + lane] : 0.0;
        // RHS at u: rates of progress stay in sh.q
        double Ti, ek, rc;
        const double f = W::rhs(L, sh, lane, st * dts, rt, us, 0.0, rhoe, 0.0, &Ti, &ek, &rc);
        double qj[2];
#pragma unroll
        for (int rr = 0; rr < 2; ++rr) qj[rr] = (rr * 64 + lane < NR) ? sh.q[rr * 64 + lane] : 0.0;
        const double r = -((t0 + t1) * us + t2 * nm1 + t3 * nm2) / dts + fs * f;
        const double sc = L.sp ? us * L.imw * 1.0e6 : 0.0;
        const double Tv = wave::rdlane(us, NS);
        // exact (constant-volume) core at the energy-inverted T for Jtrue
        AC::template core<0>(L, sh, lane, sc, Ti, sh.Jx, sh.wdx, false);
        // the BDF Jacobian at the integrated T
        if (rt == H_REACTOR)
          AC::template core<1>(L, sh, lane, sc, Tv, sh.J, sh.wd, true);
        else
          AC::template core<0>(L, sh, lane, sc, Tv, sh.J, sh.wd, true);
        adj_factor<T>(L, sh, lane, false, t0 / dts, f0, LU, kstep, pivlane, dinv);
        const double D = wave_lu_solve<n, double>(LU, lane, kstep, pivlane, dinv, act ? r : 0.0);
        adj_factor<T>(L, sh, lane, true, t0 / dts, f0, LU, kstep, pivlane, dinv);
        double mu[3];
#pragma unroll
        for (int o = 0; o < 3; ++o) mu[o] = wave_lu_solve<n, double>(LU, lane, kstep, pivlane, dinv, act ? lo[o] : 0.0);
        // broadcast mu and D; contraction weights
        if (act) {
#pragma unroll
          for (int o = 0; o < 3; ++o) sh.mu[o][lane] = mu[o];
          sh.dl[lane] = D;
        }
        // thermo at the energy-inverted T for Jtrue / df/dlnA, and at the integrated T for the T row
        double cfi[7], cft[7];
        W::coeffs(L, sh, Ti, cfi);
        W::coeffs(L, sh, Tv, cft);
        const double er = (rt == E_REACTOR) ? 1.0 : 0.0;
        const double cpk = (TabMech<T>::cp_poly(cfi, Ti) - er) * RU * L.imw;  // de_k/dT, erg/g/K
        const double dcpk = TabMech<T>::dcp_poly(cfi, Ti) * RU * L.imw;
        const double ehT = (rt == H_REACTOR) ? TabMech<T>::h_poly(cft, Tv, 1.0 / Tv)
                                              : TabMech<T>::h_poly(cft, Tv, 1.0 / Tv) - 1.0;
        const double cRT = (rt == H_REACTOR) ? TabMech<T>::cp_poly(cft, Tv) : TabMech<T>::cp_poly(cft, Tv) - 1.0;
        const double muT[3] = {wave::rdlane(mu[0], NS), wave::rdlane(mu[1], NS), wave::rdlane(mu[2], NS)};
        const double fT = wave::rdlane(f, NS);
        // T-row constants of the core at the integrated T
        const double cmix = wave::sum(L.sp ? cRT * sc : 0.0);
        const double tmp3 = Tv / cmix, cmixinv = 1.0 / cmix;
        const double dcmixdT = wave::sum(L.sp ? TabMech<T>::dcp_poly(cft, Tv) * sc : 0.0);
        const double ehmix = wave::sum(L.sp ? ehT * sh.wd[lane] : 0.0);
        const double tmp2 = ehmix * cmixinv * tmp3;
        (void)tmp2;
        const double bk = L.sp ? D / L.mw : 0.0;  // D_k / W_k
        const double ScR = wave::sum(L.sp ? cRT * 1.0e6 * bk : 0.0);
        const double DT = wave::rdlane(D, NS);
        if (L.sp) {
          sh.wb[lane] = bk;
#pragma unroll
          for (int o = 0; o < 3; ++o) {
            sh.wa[o][lane] = mu[o] * L.mw - 1.0e6 * tmp3 * ehT * muT[o];
            sh.wc[o][lane] = muT[o] * (ehT * cmixinv * tmp3 * ScR +
                                       DT * (-ehT * cmixinv + ehT * cmixinv * tmp3 * dcmixdT -
                                             tmp3 / Tv * (cRT - ehT)));
            sh.wf[o][lane] = L.mw * 1.0e-6 * (mu[o] - muT[o] * ek / rc);
          }
        }
        __syncthreads();
        // parameter gradients
        double C[2][3];
        if (rt == H_REACTOR)
          AC::template reaction_contract<1>(sh, lane, Tv, C);
        else
          AC::template reaction_contract<0>(sh, lane, Tv, C);
#pragma unroll
        for (int rr = 0; rr < 2; ++rr) {
          const int j = rr * 64 + lane;
          if (rr * 64 < NR && j < NR) {
            double dfm[3] = {0.0, 0.0, 0.0};
            for (int i = 0; i < T::MAXNET; ++i) {
              const int m = T::rxNet_(T::MAXNET * j + i);
              if (m < 0) break;
              const double nu = T::rxNu_(T::MAXNET * j + i);
#pragma unroll
              for (int o = 0; o < 3; ++o) dfm[o] += nu * sh.wf[o][m];
            }
#pragma unroll
            for (int o = 0; o < 3; ++o) g[rr][o] += fs * qj[rr] * dfm[o] + f0 * C[rr][o];
          }
        }
        // Jtrue^T mu (species components; df/dT_integrated = 0)
        const double dTdy = L.sp ? -ek / rc : 0.0;
        const double Fs = wave::sum(L.sp ? f * cpk : 0.0);
        const double Ydcp = wave::sum(L.sp ? us * dcpk : 0.0);
        double jtm[3];
#pragma unroll
        for (int o = 0; o < 3; ++o) {
          const double nuk = L.sp ? (mu[o] - muT[o] * ek / rc) * L.mw : 0.0;  // nu_k W_k
          const double sT = wave::sum(L.sp ? nuk * sh.Jx[NS * n + lane] : 0.0);
          if (L.sp) sh.wa[o][lane] = nuk;  // reuse as broadcast of nu_k W_k
          jtm[o] = sT;
        }
        __syncthreads();
        double JTmu[3] = {0.0, 0.0, 0.0};
        if (L.sp) {
          double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll 1
          for (int k = 0; k < NS; ++k) {
            const double Jv = sh.Jx[lane * n + k];
            s0 += sh.wa[0][k] * Jv;
            s1 += sh.wa[1][k] * Jv;
            s2 += sh.wa[2][k] * Jv;
          }
          const double drc = cpk + Ydcp * dTdy;
          const double ss[3] = {s0, s1, s2};
#pragma unroll
          for (int o = 0; o < 3; ++o)
            JTmu[o] = ss[o] * L.imw + dTdy * 1.0e-6 * jtm[o] - muT[o] * (Fs * dTdy + fT * drc) / rc;
        }
        __syncthreads();
        // second-order term: grad_u(mu^T J(u) D), one dual pass per component
        double Hl[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
        for (int l = 0; l < n; ++l) {
          double H[3];
          if (rt == H_REACTOR)
            AC::template dual_contract<1>(L, sh, lane, sc, Tv, l, H);
          else
            AC::template dual_contract<0>(L, sh, lane, sc, Tv, l, H);
          if (lane == l) {
            Hl[0] = H[0];
            Hl[1] = H[1];
            Hl[2] = H[2];
          }
        }
        // adjoint recursion
        const int s1 = st - (st >= 1 ? st - 1 : 0), s2 = st - (st >= 2 ? st - 2 : 0);
#pragma unroll
        for (int o = 0; o < 3; ++o) {
          const double add0 = lo[o] - (t0 + t1) / dts * mu[o] + fs * JTmu[o] + f0 * Hl[o];
          const double add1 = -t2 / dts * mu[o], add2 = -t3 / dts * mu[o];
          double acc[3] = {lam[o][0], lam[o][1], lam[o][2]};
          // slots are relative to u_st: slot 0 = u_st, 1 = u_{st-1}, 2 = u_{st-2}
          // (after the shift above lam[o][0] holds the pending adjoint of u_st)
          acc[0] += add0;
          acc[s1] += add1;
          acc[s2] += add2;
          lam[o][0] = act ? acc[0] : 0.0;
          lam[o][1] = act ? acc[1] : 0.0;
          lam[o][2] = act ? acc[2] : 0.0;
        }
      }
    }
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
      const int j = rr * 64 + lane;
      if (rr * 64 < NR && j < NR) {
        const int jm = T::rxIndex_(j);
#pragma unroll
        for (int o = 0; o < 3; ++o) a.grad[((long long)cell * NR + jm) * 3 + o] = g[rr][o];
      }
    }
    __syncthreads();
  }
}

template <class T>
__global__ void __launch_bounds__(64) wave_adjoint_kernel(ReactParamsDev P, AdjointArgs a) {
  __shared__ typename AdjChem<T>::Shared sh;
  const int lane = wave::lane_id();
  typename WaveChem<T>::Lane L;
  WaveChem<T>::load_lane(L, sh, lane);
  for (int cell = blockIdx.x; cell < a.ncells; cell += gridDim.x) wave_adjoint_cell<T>(P, a, cell, sh, L, lane);
}

}  // namespace ppc
