// Table-driven mechanism evaluation for one lane = one cell.
//
// TabMech<T> exposes the interface of the generated unrolled mechanism
// structs (mech/<name>.h) -- species thermo, mixture h/cp, production rates,
// progress rates and the analytic Jacobian with the constant-pressure
// convention -- but walks the constant-memory tables of tab/<name>.h with a
// loop index that is the same for every lane of the wavefront: parameters
// arrive through scalar loads, the loop body is shared by all reactions of a
// group, and the instruction footprint stays small enough for the
// instruction cache (the unrolled form of drm19 is ~0.5 MB of code).
//
// Equilibrium constants use per-species exponentials: with a_k = exp(-g_k/RT)
// and b_k = c_k exp(g_k/RT), the reverse rate of reaction j is
// k_f refC^delta prod_reactants a_k prod_products b_k, the same quantity as
// k_f exp(-(sum nu' g - sum nu'' g)) refC^delta prod c'' of the reference
// (Mechanisms/drm19/mechanism.H:4421-5667) with NS instead of NR exponentials;
// partial products stay below e^470 for 200 K <= T <= 4000 K.
//
// QSS mechanisms (T::NQ > 0) carry NQ species in quasi-steady state after the
// NS transported ones.  Every rate evaluation first solves their balances (a
// sparse NQ x NQ system assembled from the rates of the QSS reactions with
// the QSS concentration factored out, eliminated by the program the table
// generator wrote), then evaluates all reactions with those concentrations.
#pragma once
#include "ppc_common.h"
#include "dual.h"

namespace ppc {

// regularisation of the QSS balances (comp_sc_qss epsilon of the reference's
// generated QSS mechanisms)
constexpr double QSS_EPS = 1e-12;

template <class T>
struct TabMech {
  static constexpr int NS = T::NS, NR = T::NR, NEQ = T::NEQ;
  static constexpr int NQ = T::NQ, NX = T::NX;
  static constexpr bool kTable = true;

  PPC_DI static double mw(int k) { return T::mw_(k); }
  PPC_DI static double imw(int k) { return T::imw_(k); }
  PPC_DI static double mw_rt(int k) { return T::mw_(k); }
  PPC_DI static double imw_rt(int k) { return T::imw_(k); }

  // NASA7 coefficients of species k for temperature Tv (per-lane range select)
  PPC_DI static void nasa(int k, double Tv, double* a) {
    const bool lo = Tv < T::tmid_(k);
#pragma unroll
    for (int i = 0; i < 7; ++i) a[i] = lo ? T::nasaLo_(7 * k + i) : T::nasaHi_(7 * k + i);
  }
  // S: double, or Dual for derivatives along T
  template <class S>
  PPC_DI static S cp_poly(const double* a, S Tv) {
    return a[0] + Tv * (a[1] + Tv * (a[2] + Tv * (a[3] + Tv * a[4])));
  }
  template <class S>
  PPC_DI static S h_poly(const double* a, S Tv, S invT) {
    return a[0] + Tv * (a[1] * 0.5 + Tv * (a[2] * (1.0 / 3.0) + Tv * (a[3] * 0.25 + Tv * a[4] * 0.2))) +
           a[5] * invT;
  }
  template <class S>
  PPC_DI static S g_poly(const double* a, S Tv, S invT, S logT) {
    return a[5] * invT + (a[0] - a[6]) - a[0] * logT -
           Tv * (a[1] * 0.5 + Tv * (a[2] * (1.0 / 6.0) + Tv * (a[3] * (1.0 / 12.0) + Tv * a[4] * 0.05)));
  }
  // d(cp/R)/dT
  template <class S>
  PPC_DI static S dcp_poly(const double* a, S Tv) {
    return a[1] + Tv * (2.0 * a[2] + Tv * (3.0 * a[3] + Tv * 4.0 * a[4]));
  }

  PPC_DI static void cp_R(double* out, double Tv) {
#pragma unroll 1
    for (int k = 0; k < NS; ++k) {
      double a[7];
      nasa(k, Tv, a);
      out[k] = cp_poly(a, Tv);
    }
  }
  PPC_DI static void cv_R(double* out, double Tv) {
    cp_R(out, Tv);
    for (int k = 0; k < NS; ++k) out[k] -= 1.0;
  }
  PPC_DI static void h_RT(double* out, double Tv) {
    const double invT = 1.0 / Tv;
#pragma unroll 1
    for (int k = 0; k < NS; ++k) {
      double a[7];
      nasa(k, Tv, a);
      out[k] = h_poly(a, Tv, invT);
    }
  }
  PPC_DI static void e_RT(double* out, double Tv) {
    h_RT(out, Tv);
    for (int k = 0; k < NS; ++k) out[k] -= 1.0;
  }
  PPC_DI static void gibbs_RT(double* out, double Tv) {
    const double invT = 1.0 / Tv, logT = log(Tv);
#pragma unroll 1
    for (int k = 0; k < NS; ++k) {
      double a[7];
      nasa(k, Tv, a);
      out[k] = g_poly(a, Tv, invT, logT);
    }
  }
  PPC_DI static void dcpR_dT(double* out, double Tv) {
#pragma unroll 1
    for (int k = 0; k < NS; ++k) {
      double a[7];
      nasa(k, Tv, a);
      out[k] = a[1] + Tv * (2.0 * a[2] + Tv * (3.0 * a[3] + Tv * 4.0 * a[4]));
    }
  }
  // mixture h (erg/g) and cp (erg/g/K), or e and cv, from mass fractions
  PPC_DI static void h_cp_mass(double Tv, const double* Y, double& hm, double& cm) {
    const double invT = 1.0 / Tv;
    double hs = 0.0, cs = 0.0;
#pragma unroll 1
    for (int k = 0; k < NS; ++k) {
      double a[7];
      nasa(k, Tv, a);
      const double yw = Y[k] * T::imw_(k);
      hs += yw * h_poly(a, Tv, invT);
      cs += yw * cp_poly(a, Tv);
    }
    hm = hs * (RU * Tv);
    cm = cs * RU;
  }
  PPC_DI static void e_cv_mass(double Tv, const double* Y, double& em, double& cm) {
    const double invT = 1.0 / Tv;
    double es = 0.0, cs = 0.0;
#pragma unroll 1
    for (int k = 0; k < NS; ++k) {
      double a[7];
      nasa(k, Tv, a);
      const double yw = Y[k] * T::imw_(k);
      es += yw * (h_poly(a, Tv, invT) - 1.0);
      cs += yw * (cp_poly(a, Tv) - 1.0);
    }
    em = es * (RU * Tv);
    cm = cs * RU;
  }

  // refC^delta for the small integer powers that occur
  template <class S>
  PPC_DI static S refc_pow(int d, S refC, S refCinv) {
    S r = 1.0;
    for (int i = 0; i < d; ++i) r *= refC;
    for (int i = 0; i < -d; ++i) r *= refCinv;
    return r;
  }

  // Net rate of progress of reaction j (table order) given c (sc), a = exp(-g),
  // b = c exp(g); returns q and (optionally) the pieces the Jacobian needs.
  template <int KIND>  // 0 elementary, 1 three-body, 2 Troe falloff, 3 Lindemann falloff
  PPC_DI static double rate(int j, int fo, const double* sc, const double* a, const double* b, double mixture,
                            double Tv, double invT, double logT, double refC, double refCinv) {
    double qf, qr, dlnc, kfc;
    rate2<KIND, false>(j, fo, sc, a, b, mixture, Tv, invT, logT, refC, refCinv, qf, qr, dlnc, kfc);
    return qf - qr;
  }
  // The forward and reverse rates of reaction j (the forward rate 0 where the
  // QSS reduction removed it); with DLNC also d ln(corr) / d alpha of the
  // third-body / falloff factor corr (alpha: the enhanced concentration) and
  // kfc = corr k_f.
  template <int KIND, bool DLNC>
  PPC_DI static void rate2(int j, int fo, const double* sc, const double* a, const double* b, double mixture,
                           double Tv, double invT, double logT, double refC, double refCinv, double& qf_o,
                           double& qr_o, double& dlnc, double& kfc) {
    const double kf0 = T::rxA_(j) * exp(T::rxB_(j) * logT - T::rxE_(j) * invT);
    double corr = 1.0;
    dlnc = 0.0;
    if (KIND >= 1) {
      double alpha = mixture;
      const int e0 = T::rxEff0_(j), en = T::rxEffN_(j);
#pragma unroll 1
      for (int i = 0; i < en; ++i) alpha += T::effV_(e0 + i) * sc[T::effK_(e0 + i)];
      if (KIND == 1) {
        corr = alpha;
        if (DLNC) dlnc = 1.0 / alpha;
      } else {
        const double k0 = T::foA_(fo) * exp(T::foB_(fo) * logT - T::foE_(fo) * invT);
        const double redP = alpha / kf0 * k0;
        const double F = redP / (1.0 + redP);
        if (KIND == 2) {
          const double tA = T::troe_(4 * fo), t3 = T::troe_(4 * fo + 1), t1 = T::troe_(4 * fo + 2);
          double Fcent = (1.0 - tA) * exp(-Tv / t3) + tA * exp(-Tv / t1);
          if (T::troeT2_(fo)) Fcent += exp(-T::troe_(4 * fo + 3) * invT);
          const double logFcent = log10(Fcent);
          const double logPred = log10(redP);
          const double tc = -0.4 - 0.67 * logFcent;
          const double tn = 0.75 - 1.27 * logFcent;
          const double tr = (tc + logPred) / (tn - 0.14 * (tc + logPred));
          corr = F * exp(LN10 * logFcent / (1.0 + tr * tr));
          if (DLNC) {
            // d log10 F / d log10 Pr of the Troe form, as in jac_reaction
            const double den = 1.0 / (tn - 0.14 * (tc + logPred));
            const double tPr = (tc + logPred) * den;
            const double troe = 1.0 / (1.0 + tPr * tPr);
            dlnc = (1.0 / (1.0 + redP) - tn * 2.0 * logFcent * troe * troe * tPr * den * den) / alpha;
          }
        } else {
          corr = F;
          if (DLNC) dlnc = 1.0 / ((1.0 + redP) * alpha);
        }
      }
    }
    const double kf = corr * kf0;
    double qf = kf;
    double qr = 0.0;
    const int r0 = T::rxR_(3 * j), r1 = T::rxR_(3 * j + 1), r2 = T::rxR_(3 * j + 2);
    qf *= sc[r0];
    if (r1 >= 0) qf *= sc[r1];
    if (r2 >= 0) qf *= sc[r2];
    if constexpr (NQ > 0)
      if (T::rxFwdOff_(j)) qf = 0.0;
    if (T::rxRev_(j)) {
      const int p0 = T::rxP_(3 * j), p1 = T::rxP_(3 * j + 1), p2 = T::rxP_(3 * j + 2);
      qr = kf * refc_pow(T::rxDelta_(j), refC, refCinv) * a[r0] * b[p0];
      if (r1 >= 0) qr *= a[r1];
      if (r2 >= 0) qr *= a[r2];
      if (p1 >= 0) qr *= b[p1];
      if (p2 >= 0) qr *= b[p2];
    }
    qf_o = qf;
    qr_o = qr;
    kfc = kf;
  }

  // rate2 of reaction j (table order) with its group's kind
  template <bool DLNC>
  PPC_DI static void rate_any(int j, const double* sc, const double* a, const double* b, double mixture, double Tv,
                              double invT, double logT, double refC, double refCinv, double& qf, double& qr,
                              double& dlnc, double& kfc) {
    const int fo = j - T::N_EL - T::N_TB;
    if (j < T::N_EL)
      rate2<0, DLNC>(j, 0, sc, a, b, mixture, Tv, invT, logT, refC, refCinv, qf, qr, dlnc, kfc);
    else if (j < T::N_EL + T::N_TB)
      rate2<1, DLNC>(j, 0, sc, a, b, mixture, Tv, invT, logT, refC, refCinv, qf, qr, dlnc, kfc);
    else if (j < NR - T::N_LIND)
      rate2<2, DLNC>(j, fo, sc, a, b, mixture, Tv, invT, logT, refC, refCinv, qf, qr, dlnc, kfc);
    else
      rate2<3, DLNC>(j, fo, sc, a, b, mixture, Tv, invT, logT, refC, refCinv, qf, qr, dlnc, kfc);
  }

  // QSS balances: with sce = (transported concentrations, ones) and a, b of
  // species_exp at those ones, assemble the QSS matrix A (T::NSLOT entries)
  // and the right-hand side r (NQ), eliminate, and put the QSS
  // concentrations x into r, sce[NS..] and b[NS..] (b_q = x_q exp(g_q)).  A
  // keeps its LU factors.
  PPC_DI static void qss_solve(double* sce, const double* a, double* b, double mixture, double Tv, double invT,
                               double logT, double refC, double refCinv, double* A, double* r) {
#pragma unroll 1
    for (int i = 0; i < T::NSLOT; ++i) A[i] = T::slotInit_(i);
    for (int q = 0; q < NQ; ++q) r[q] = QSS_EPS;
#pragma unroll 1
    for (int i = 0; i < T::NQR; ++i) {
      const int j = T::qRx_(i);
      double qf, qr, dl, kfc;
      rate_any<false>(j, sce, a, b, mixture, Tv, invT, logT, refC, refCinv, qf, qr, dl, kfc);
#pragma unroll 1
      for (int t = T::qTermPtr_(i); t < T::qTermPtr_(i + 1); ++t) {
        const double v = T::tMul_(t) * (T::tDir_(t) ? qr : qf);
        const int sl = T::tSlot_(t);
        if (sl >= 0)
          A[sl] += v;
        else
          r[-1 - sl] += v;
      }
    }
    lu_qss(A);
    solve_qss(A, r);
    for (int q = 0; q < NQ; ++q) {
      sce[NS + q] = r[q];
      b[NS + q] *= r[q];
    }
  }

  PPC_DI static void lu_qss(double* A) {
#pragma unroll 1
    for (int i = 0; i < T::NLU; ++i) {
      const int c = T::luC_(i);
      if (c < 0)
        A[T::luA_(i)] /= A[T::luB_(i)];
      else
        A[T::luA_(i)] -= A[T::luB_(i)] * A[c];
    }
  }
  // solves A x = r in place with the factors of lu_qss
  PPC_DI static void solve_qss(const double* A, double* r) {
#pragma unroll 1
    for (int i = 0; i < T::NFS; ++i) r[T::fsI_(i)] -= A[T::fsS_(i)] * r[T::fsP_(i)];
#pragma unroll 1
    for (int i = 0; i < NQ; ++i) {
      const int row = T::bsRow_(i);
      double acc = r[row];
#pragma unroll 1
      for (int e = T::bsPtr_(i); e < T::bsPtr_(i + 1); ++e) acc -= A[T::bsSlot_(e)] * r[T::bsCol_(e)];
      r[row] = acc / A[T::bsDiag_(i)];
    }
  }

  // a_k = exp(-g_k/RT), b_k = c_k exp(g_k/RT) over all NX species (QSS
  // species included), mixture = sum of the transported c
  PPC_DI static double species_exp(const double* sc, double Tv, double* a, double* b) {
    const double invT = 1.0 / Tv, logT = log(Tv);
    double mixture = 0.0;
#pragma unroll 1
    for (int k = 0; k < NX; ++k) {
      double c[7];
      nasa(k, Tv, c);
      const double eg = exp(g_poly(c, Tv, invT, logT));
      a[k] = 1.0 / eg;
      b[k] = sc[k] * eg;
      if (k < NS) mixture += sc[k];
    }
    return mixture;
  }

  // sce = (sc, QSS concentrations) with a, b of species_exp at them; A and r
  // hold the QSS factors and concentrations (QSS mechanisms only)
  PPC_DI static double qss_state(const double* sc, double Tv, double* sce, double* a, double* b, double* A,
                                 double* r) {
    for (int k = 0; k < NS; ++k) sce[k] = sc[k];
    for (int k = NS; k < NX; ++k) sce[k] = 1.0;
    const double mixture = species_exp(sce, Tv, a, b);
    const double invT = 1.0 / Tv, logT = log(Tv);
    const double refC = 101325.0 / 8.31446 * invT, refCinv = 1.0 / refC;
    qss_solve(sce, a, b, mixture, Tv, invT, logT, refC, refCinv, A, r);
    return mixture;
  }

  template <bool WDOT>
  PPC_DI static void rates(double* __restrict__ out, const double* __restrict__ sc_in, const double Tv) {
    double a[NX], b[NX];
    double sce[NQ > 0 ? NX : 1], Aq[T::NSLOT], rq[NQ > 0 ? NQ : 1];
    const double* sc = sc_in;
    double mixture;
    if constexpr (NQ > 0) {
      mixture = qss_state(sc_in, Tv, sce, a, b, Aq, rq);
      sc = sce;
    } else {
      mixture = species_exp(sc_in, Tv, a, b);
    }
    const double invT = 1.0 / Tv, logT = log(Tv);
    const double refC = 101325.0 / 8.31446 * invT, refCinv = 1.0 / refC;
    if (WDOT)
      for (int k = 0; k < NS; ++k) out[k] = 0.0;
    auto put = [&](int j, double q) {
      if (WDOT) {
        for (int i = 0; i < T::MAXNET; ++i) {
          const int k = T::rxNet_(T::MAXNET * j + i);
          if (k < 0) break;
          out[k] += T::rxNu_(T::MAXNET * j + i) * q;
        }
      } else {
        out[T::rxIndex_(j)] = q;
      }
    };
    int j = 0;
#pragma unroll 1
    for (; j < T::N_EL; ++j) put(j, rate<0>(j, 0, sc, a, b, mixture, Tv, invT, logT, refC, refCinv));
#pragma unroll 1
    for (; j < T::N_EL + T::N_TB; ++j) put(j, rate<1>(j, 0, sc, a, b, mixture, Tv, invT, logT, refC, refCinv));
#pragma unroll 1
    for (int f = 0; f < T::N_TROE; ++f, ++j)
      put(j, rate<2>(j, f, sc, a, b, mixture, Tv, invT, logT, refC, refCinv));
#pragma unroll 1
    for (int f = T::N_TROE; f < T::N_TROE + T::N_LIND; ++f, ++j)
      put(j, rate<3>(j, f, sc, a, b, mixture, Tv, invT, logT, refC, refCinv));
  }

  // molar production rates (mol/m^3/s) from concentrations (mol/m^3)
  PPC_DI static void production_rate(double* wdot, const double* sc, double Tv) { rates<true>(wdot, sc, Tv); }
  // net rates of progress in YAML order
  PPC_DI static void progress_rates(double* q, const double* sc, double Tv) { rates<false>(q, sc, Tv); }

  // Contribution of reaction j (table order, falloff index fo) to the
  // analytic Jacobian: addJ(idx, v) adds v to J[idx] (column-major,
  // idx = k*(NS+1)+m for d wdot_m / d c_k, NS*(NS+1)+m for d wdot_m / dT) and
  // addW(m, v) to wdot_m.  g and hRT are the species g/RT and h/RT at Tv.
  // S is double, or Dual for directional derivatives of the whole contribution.
  template <int CONSP, class S, class AddJ, class AddW>
  PPC_DI static void jac_reaction(int j, const S* __restrict__ sc, const S* __restrict__ g,
                                  const S* __restrict__ hRT, S mixture, S Tv, S invT, S invT2, S logT, S refC,
                                  S refCinv, AddJ&& addJ, AddW&& addW) {
    constexpr int n = NS + 1;
    const int kind = j < T::N_EL ? 0 : (j < T::N_EL + T::N_TB ? 1 : (j < NR - T::N_LIND ? 2 : 3));
    const int fo = j - T::N_EL - T::N_TB;
    const int r0 = T::rxR_(3 * j), r1 = T::rxR_(3 * j + 1), r2 = T::rxR_(3 * j + 2);
    const int p0 = T::rxP_(3 * j), p1 = T::rxP_(3 * j + 1), p2 = T::rxP_(3 * j + 2);
    const bool rev = T::rxRev_(j);
    S phi_f = sc[r0];
    if (r1 >= 0) phi_f *= sc[r1];
    if (r2 >= 0) phi_f *= sc[r2];
    S k_f = T::rxA_(j) * exp(T::rxB_(j) * logT - T::rxE_(j) * invT);
    const S dlnkfdT = T::rxB_(j) * invT + T::rxE_(j) * invT2;
    S alpha = 0.0, Pr = 0.0, fPr = 1.0, F = 1.0, dlogFdlogPr = 0.0, dlogFdT = 0.0, dlogfPrdT = 0.0;
    if (kind >= 1) {
      alpha = mixture;
      const int e0 = T::rxEff0_(j), en = T::rxEffN_(j);
#pragma unroll 1
      for (int i = 0; i < en; ++i) alpha += T::effV_(e0 + i) * sc[T::effK_(e0 + i)];
    }
    if (kind >= 2) {
      const S k0 = T::foA_(fo) * exp(T::foB_(fo) * logT - T::foE_(fo) * invT);
      Pr = alpha / k_f * k0;
      fPr = Pr / (1.0 + Pr);
      const S dlnk0dT = T::foB_(fo) * invT + T::foE_(fo) * invT2;
      const S dlogPrdT = LOG10E * (dlnk0dT - dlnkfdT);
      dlogfPrdT = dlogPrdT / (1.0 + Pr);
      if (kind == 2) {
        const double tA = T::troe_(4 * fo), t3 = T::troe_(4 * fo + 1), t1 = T::troe_(4 * fo + 2);
        const double t2 = T::troe_(4 * fo + 3);
        const bool has2 = T::troeT2_(fo);
        const S logPr = log10(Pr);
        const S F1 = (1.0 - tA) * exp(-Tv / t3);
        const S F2 = tA * exp(-Tv / t1);
        const S F3 = has2 ? exp(-t2 * invT) : S(0.0);
        const S Fcent = F1 + F2 + F3;
        const S logFcent = log10(Fcent);
        const S tc = -0.4 - 0.67 * logFcent, tn = 0.75 - 1.27 * logFcent;
        const S den = 1.0 / (tn - 0.14 * (tc + logPr));
        const S tPr = (tc + logPr) * den;
        const S troe = 1.0 / (1.0 + tPr * tPr);
        F = exp(LN10 * logFcent * troe);
        const S dlogFcentdT = LOG10E / Fcent * (-F1 / t3 - F2 / t1 + (has2 ? F3 * t2 * invT2 : S(0.0)));
        const S fac = 2.0 * logFcent * troe * troe * tPr * den;
        const S dlogFdc = -tn * fac * den;
        const S dlogFdn = fac * tPr;
        dlogFdlogPr = dlogFdc;
        dlogFdT = dlogFcentdT * (troe - 0.67 * dlogFdc - 1.27 * dlogFdn) + dlogFdlogPr * dlogPrdT;
      }
    }
    S phi_r = 0.0, k_r = 0.0, dkrdT = 0.0;
    if (rev) {
      phi_r = sc[p0];
      if (p1 >= 0) phi_r *= sc[p1];
      if (p2 >= 0) phi_r *= sc[p2];
      S dg = g[r0] - g[p0], hf = hRT[r0], hr = hRT[p0];
      if (r1 >= 0) { dg += g[r1]; hf += hRT[r1]; }
      if (r2 >= 0) { dg += g[r2]; hf += hRT[r2]; }
      if (p1 >= 0) { dg -= g[p1]; hr += hRT[p1]; }
      if (p2 >= 0) { dg -= g[p2]; hr += hRT[p2]; }
      const int d = T::rxDelta_(j);
      k_r = k_f * exp(-dg) * refc_pow(d, refC, refCinv);
      const S dlnKcdT = invT * (-hf + hr + (double)d);
      dkrdT = (dlnkfdT - dlnKcdT) * k_r;
    }
    bool fwd = true;  // forward direction present (not removed by the QSS reduction)
    if constexpr (NQ > 0) {
      if (T::rxFwdOff_(j)) {
        fwd = false;
        phi_f = 0.0;
      }
    }
    S q, dqdT, fac = 0.0;
    if (kind == 0) {
      q = k_f * phi_f - k_r * phi_r;
      dqdT = dlnkfdT * k_f * phi_f - dkrdT * phi_r;
    } else if (kind == 1) {
      const S qn = k_f * phi_f - k_r * phi_r;
      q = alpha * qn;
      dqdT = alpha * (dlnkfdT * k_f * phi_f - dkrdT * phi_r);
      k_f *= alpha;
      k_r *= alpha;
      fac = qn;
    } else {
      const S qn = k_f * phi_f - k_r * phi_r;
      const S Corr = fPr * F;
      q = Corr * qn;
      const S dlnCorrdT = LN10 * (dlogfPrdT + dlogFdT);
      dqdT = Corr * (dlnkfdT * k_f * phi_f - dkrdT * phi_r) + dlnCorrdT * q;
      k_f *= Corr;
      k_r *= Corr;
      fac = q / alpha * (1.0 / (Pr + 1.0) + dlogFdlogPr);
    }
    // net stoichiometry of this reaction
    int nk[T::MAXNET];
    double nv[T::MAXNET];
    int nn = 0;
    for (int i = 0; i < T::MAXNET; ++i) {
      nk[i] = T::rxNet_(T::MAXNET * j + i);
      nv[i] = T::rxNu_(T::MAXNET * j + i);
      if (nk[i] >= 0) nn = i + 1;
    }
    for (int i = 0; i < nn; ++i) {
      addW(nk[i], nv[i] * q);
      addJ(NS * n + nk[i], nv[i] * dqdT);
    }
    auto col = [&](int k, const S& dq) {
      if (NQ > 0 && k >= NS) return;  // QSS columns: handled through the QSS balances
      for (int i = 0; i < nn; ++i) addJ(k * n + nk[i], nv[i] * dq);
    };
    // d(prod c^nu)/dc_k over the expanded species lists (repeated index = coefficient 2, 3)
    const int rl[3] = {r0, r1, r2}, pl[3] = {p0, p1, p2};
    for (int i = 0; i < 3; ++i) {
      const int k = rl[i];
      if (!fwd || k < 0 || (i > 0 && k == rl[i - 1])) continue;
      int mult = 0;
      S rest = 1.0;
      for (int l = 0; l < 3; ++l) {
        if (rl[l] < 0) continue;
        if (rl[l] == k) {
          if (mult > 0) rest *= sc[k];
          ++mult;
        } else {
          rest *= sc[rl[l]];
        }
      }
      col(k, k_f * (double)mult * rest);
    }
    if (rev) {
      for (int i = 0; i < 3; ++i) {
        const int k = pl[i];
        if (k < 0 || (i > 0 && k == pl[i - 1])) continue;
        int mult = 0;
        S rest = 1.0;
        for (int l = 0; l < 3; ++l) {
          if (pl[l] < 0) continue;
          if (pl[l] == k) {
            if (mult > 0) rest *= sc[k];
            ++mult;
          } else {
            rest *= sc[pl[l]];
          }
        }
        col(k, -k_r * (double)mult * rest);
      }
    }
    if (kind >= 1) {
      if (CONSP == 0) {
#pragma unroll 1
        for (int k = 0; k < NS; ++k) col(k, fac);
      }
      const int e0 = T::rxEff0_(j), en = T::rxEffN_(j);
#pragma unroll 1
      for (int i = 0; i < en; ++i) col(T::effK_(e0 + i), T::effV_(e0 + i) * fac);
    }
  }

  // Analytic Jacobian in SI-mol units, column-major J[k*(NS+1)+m] = d wdot_m / d c_k,
  // J[NS*(NS+1)+m] = d wdot_m / dT, T row d Tdot / d c_k, d Tdot / dT
  // (aJacobian semantics, Mechanisms/LiDryer/mechanism.H:3905-5138).
  // J is a double* or a LaneArr<double> (lane_ws.h).
  template <int CONSP, class JM>
  PPC_DI static void jacobian(const JM J, const double* __restrict__ sc, const double Tv) {
    if constexpr (NQ > 0) {
      jacobian_qss<CONSP>(J, sc, Tv);
      return;
    }
    constexpr int n = NS + 1;
#pragma unroll 1
    for (int i = 0; i < n * n; ++i) J[i] = 0.0;
    double wdot[NS], hRT[NS], g[NS];
    for (int k = 0; k < NS; ++k) wdot[k] = 0.0;
    const double invT = 1.0 / Tv, invT2 = invT * invT, logT = log(Tv);
    const double refC = 101325.0 / 8.31446 * invT, refCinv = 1.0 / refC;
    double mixture = 0.0;
    for (int k = 0; k < NS; ++k) mixture += sc[k];
    gibbs_RT(g, Tv);
    h_RT(hRT, Tv);
#pragma unroll 1
    for (int j = 0; j < NR; ++j)
      jac_reaction<CONSP>(j, sc, g, hRT, mixture, Tv, invT, invT2, logT, refC, refCinv,
                          [&](int idx, double v) { J[idx] += v; }, [&](int m, double v) { wdot[m] += v; });
    // temperature row (mechanism.H:5098-5137)
    double cR[NS], dcR[NS];
    cp_R(cR, Tv);
    if (CONSP != 1)
      for (int k = 0; k < NS; ++k) cR[k] -= 1.0;
    dcpR_dT(dcR, Tv);
    double cmix = 0.0, ehmix = 0.0, dcmixdT = 0.0, dehmixdT = 0.0;
#pragma unroll 1
    for (int k = 0; k < NS; ++k) {
      const double eh = (CONSP == 1) ? hRT[k] : hRT[k] - 1.0;
      cmix += cR[k] * sc[k];
      dcmixdT += dcR[k] * sc[k];
      ehmix += eh * wdot[k];
      dehmixdT += invT * (cR[k] - eh) * wdot[k] + eh * J[NS * n + k];
    }
    const double cmixinv = 1.0 / cmix;
    const double tmp1 = ehmix * cmixinv, tmp3 = cmixinv * Tv, tmp2 = tmp1 * tmp3;
#pragma unroll 1
    for (int k = 0; k < NS; ++k) {
      double d = 0.0;
#pragma unroll 1
      for (int mm = 0; mm < NS; ++mm) d += ((CONSP == 1) ? hRT[mm] : hRT[mm] - 1.0) * J[k * n + mm];
      J[k * n + NS] = tmp2 * cR[k] - tmp3 * d;
    }
    J[n * n - 1] = -tmp1 + tmp2 * dcmixdT - tmp3 * dehmixdT;
  }

  // QSS mechanisms: the species block is the exact derivative of the
  // QSS-reduced production rates (the reference differentiates the same
  // expressions symbolically, Support/ceptr/ceptr/jacobian.py ajac_symbolic):
  // the direct part at fixed QSS concentrations x, with third-body sums over
  // every species whatever CONSP, plus (d wdot / d x)(d x / d c) with
  // d x / d c = -A^-1 d(balance)/d c.  The temperature column is the forward
  // difference of the production rates with dT = 1e-2 K and the temperature
  // row follows from both (jacobian.py:386-470,
  // Mechanisms/CH4_lean_qss/mechanism.H:9013-9080).
  template <int CONSP, class JM>
  PPC_DI static void jacobian_qss(const JM J, const double* __restrict__ sc, const double Tv) {
    constexpr int n = NS + 1;
#pragma unroll 1
    for (int i = 0; i < n * n; ++i) J[i] = 0.0;
    double sce[NX], a[NX], b[NX], Aq[T::NSLOT], x[NQ];
    const double mixture = qss_state(sc, Tv, sce, a, b, Aq, x);
    const double invT = 1.0 / Tv, invT2 = invT * invT, logT = log(Tv);
    const double refC = 101325.0 / 8.31446 * invT, refCinv = 1.0 / refC;
    double g[NX], hRT[NX];
#pragma unroll 1
    for (int k = 0; k < NX; ++k) {
      double c[7];
      nasa(k, Tv, c);
      g[k] = g_poly(c, Tv, invT, logT);
      hRT[k] = h_poly(c, Tv, invT);
    }
#pragma unroll 1
    for (int j = 0; j < NR; ++j)
      jac_reaction<0>(j, sce, g, hRT, mixture, Tv, invT, invT2, logT, refC, refCinv,
                      [&](int idx, double v) {
                        if (idx < NS * n) J[idx] += v;
                      },
                      [&](int, double) {});
    // the transported concentrations with the QSS ones set to 1 (and b with
    // them): the rates with the QSS concentration factored out
    double sc1[NX], b1[NX];
    for (int k = 0; k < NS; ++k) {
      sc1[k] = sce[k];
      b1[k] = b[k];
    }
    for (int k = NS; k < NX; ++k) {
      sc1[k] = 1.0;
      b1[k] = exp(g[k]);
    }
    // G = d(balance)/d c, NQ x NS row-major
    double G[NQ * NS];
    for (int i = 0; i < NQ * NS; ++i) G[i] = 0.0;
#pragma unroll 1
    for (int i = 0; i < T::NQR; ++i) {
      const int j = T::qRx_(i);
      double qf, qr, dlnc, kfc;
      rate_any<true>(j, sc1, a, b1, mixture, Tv, invT, logT, refC, refCinv, qf, qr, dlnc, kfc);
      const bool tb = j >= T::N_EL;
      // k_r corr = kfc refC^delta prod a_reactants / prod a_products
      double krc = 0.0;
      if (T::rxRev_(j)) {
        krc = kfc * refc_pow(T::rxDelta_(j), refC, refCinv);
        for (int l = 0; l < 3; ++l) {
          const int r = T::rxR_(3 * j + l), p = T::rxP_(3 * j + l);
          if (r >= 0) krc *= a[r];
          if (p >= 0) krc /= a[p];
        }
      }
#pragma unroll 1
      for (int t = T::qTermPtr_(i); t < T::qTermPtr_(i + 1); ++t) {
        const int dir = T::tDir_(t), row = T::tRow_(t), cq = T::tCol_(t);
        if (dir == 0 && T::rxFwdOff_(j)) continue;
        const double w = cq < 0 ? -T::tMul_(t) : T::tMul_(t) * x[cq];
        const double qco = dir ? qr : qf;
        const double kd = dir ? krc : kfc;
        int sl[3];
        for (int l = 0; l < 3; ++l) sl[l] = dir ? T::rxP_(3 * j + l) : T::rxR_(3 * j + l);
        for (int l = 0; l < 3; ++l) {
          const int k = sl[l];
          if (k < 0 || k >= NS || (l > 0 && k == sl[l - 1])) continue;
          int mult = 0;
          double rest = 1.0;
          for (int u = 0; u < 3; ++u) {
            if (sl[u] < 0) continue;
            if (sl[u] == k) {
              if (mult > 0) rest *= sc1[k];
              ++mult;
            } else {
              rest *= sc1[sl[u]];
            }
          }
          G[row * NS + k] += w * kd * (double)mult * rest;
        }
        if (tb) {
          const double d = w * dlnc * qco;
#pragma unroll 1
          for (int k = 0; k < NS; ++k) G[row * NS + k] += d;
          const int e0 = T::rxEff0_(j), en = T::rxEffN_(j);
#pragma unroll 1
          for (int e = 0; e < en; ++e) G[row * NS + T::effK_(e0 + e)] += T::effV_(e0 + e) * d;
        }
      }
    }
    // X = -A^-1 G, column by column, in place
#pragma unroll 1
    for (int k = 0; k < NS; ++k) {
      double col[NQ];
      for (int q = 0; q < NQ; ++q) col[q] = -G[q * NS + k];
      solve_qss(Aq, col);
      for (int q = 0; q < NQ; ++q) G[q * NS + k] = col[q];
    }
    // J[:, k] += (d wdot / d x) X[:, k]: q_j is linear in its QSS reactant
    // (rate qf_co) and QSS product (rate -qr_co)
#pragma unroll 1
    for (int i = 0; i < T::NQR; ++i) {
      const int j = T::qRx_(i);
      double qf, qr, dlnc, kfc;
      rate_any<false>(j, sc1, a, b1, mixture, Tv, invT, logT, refC, refCinv, qf, qr, dlnc, kfc);
      int qR = -1, qP = -1;
      for (int l = 0; l < 3; ++l) {
        const int r = T::rxR_(3 * j + l), p = T::rxP_(3 * j + l);
        if (r >= NS) qR = r - NS;
        if (p >= NS) qP = p - NS;
      }
      if (T::rxFwdOff_(j)) qR = -1;
      if (!T::rxRev_(j)) qP = -1;
#pragma unroll 1
      for (int l = 0; l < T::MAXNET; ++l) {
        const int m = T::rxNet_(T::MAXNET * j + l);
        if (m < 0) break;
        const double nu = T::rxNu_(T::MAXNET * j + l);
#pragma unroll 1
        for (int k = 0; k < NS; ++k) {
          double d = 0.0;
          if (qR >= 0) d += qf * G[qR * NS + k];
          if (qP >= 0) d -= qr * G[qP * NS + k];
          J[k * n + m] += nu * d;
        }
      }
    }
    // temperature column (forward difference) and row
    double w0[NS], w1[NS];
    production_rate(w0, sc, Tv);
    production_rate(w1, sc, Tv + 1e-2);
    for (int k = 0; k < NS; ++k) J[NS * n + k] = (w1[k] - w0[k]) / 1e-2;
    double cR[NS], dcR[NS];
    cp_R(cR, Tv);
    if (CONSP != 1)
      for (int k = 0; k < NS; ++k) cR[k] -= 1.0;
    dcpR_dT(dcR, Tv);
    double cmix = 0.0, ehmix = 0.0, dcmixdT = 0.0, dehmixdT = 0.0;
#pragma unroll 1
    for (int k = 0; k < NS; ++k) {
      const double eh = (CONSP == 1) ? hRT[k] : hRT[k] - 1.0;
      cmix += cR[k] * sc[k];
      dcmixdT += dcR[k] * sc[k];
      ehmix += eh * w0[k];
      dehmixdT += invT * (cR[k] - eh) * w0[k] + eh * J[NS * n + k];
    }
    const double cmixinv = 1.0 / cmix;
    const double tmp1 = ehmix * cmixinv, tmp3 = cmixinv * Tv, tmp2 = tmp1 * tmp3;
#pragma unroll 1
    for (int k = 0; k < NS; ++k) {
      double d = 0.0;
#pragma unroll 1
      for (int mm = 0; mm < NS; ++mm) d += ((CONSP == 1) ? hRT[mm] : hRT[mm] - 1.0) * J[k * n + mm];
      J[k * n + NS] = tmp2 * cR[k] - tmp3 * d;
    }
    J[n * n - 1] = -tmp1 + tmp2 * dcmixdT - tmp3 * dehmixdT;
  }
};

}  // namespace ppc
