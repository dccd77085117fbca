// Batched adaptive BDF with a group of G lanes per cell (64 / G cells per
// wavefront, G = 8, 16, 32 or 64).
//
// The lane-per-cell batched kernel (batched.h) keeps a whole cell's
// integrator state in one lane: ~12 KB for drm19, which the register file
// cannot hold, so it lives in private (scratch) memory and every access waits
// on the memory system.  Here the equations of a cell are spread over the G
// lanes of its group, so the state fits the registers:
//
// * lane g of a group holds equations i = e*G + g for slots e < EPL =
//   ceil(NEQ / G): its entries of the BDF difference array, the iterate, the
//   error weights and rows i of the Newton-matrix LU factors;
// * right-hand side: each lane evaluates the NASA7 polynomials and
//   exponentials of its species, concentrations and exponentials go to a small
//   per-group LDS area, reactions j = r*G + g are spread over the lanes and
//   write their net rates of progress to LDS, and each lane sums its species'
//   production rates over the compressed species -> reaction table; mixture
//   sums (density, energy, the temperature Newton iteration) are group
//   reductions;
// * Jacobian (rare: about one per cell and call): the same reaction split,
//   accumulated with FP64 atomics into a per-group matrix in the device
//   workspace, then each lane takes its rows (molecular-weight scaling of
//   ReactorBDF.cpp:67-82) and keeps them in the workspace in LU_T for the
//   Newton-matrix refreshes;
// * Newton matrix: inverted in LU_T by row-per-lane Gauss-Jordan exchange
//   steps with partial pivoting (the pivot a group arg-max, the pivot row
//   through LDS, rows never move), so that each Newton iteration's solve is
//   one dot product per row: no sequential triangular sweeps.
//
// Every group reduction returns the bitwise-identical value in all lanes of
// the group (pairs are always added in commuted order on both partners), so
// the integrator's control flow is uniform within a group: the group's lanes
// are always active together, which the lane exchanges require.  Different
// groups of a wave follow their own cells; as in batched.h every trip of the
// persistent loop evaluates one right-hand side for every group that is not
// waiting for a factorisation, and the groups' Jacobian / LU requests are
// served together once enough of them wait.
//
// The step / order / Newton logic is that of batched_kernel (batched.h), the
// NDF family of SciPy's BDF integrator; every loop over a cell's work is
// bounded (queue length, RHS budget, the 200-iteration temperature inversion).
// The algorithm follows SciPy's scipy/integrate/_ivp/bdf.py (BSD-3-Clause,
// Copyright (c) 2001-2002 Enthought, Inc. 2003-2024, SciPy Developers; the
// notice and license are in THIRD_PARTY_NOTICES.md at the repository root).
#pragma once
#include "chem_table.h"
#include "reactors.h"
#include "wave.h"
#include "batched.h"

namespace ppc {

// ------------------------------------------------ group lane primitives --
template <int G>
struct Grp {
  static_assert(G == 8 || G == 16 || G == 32 || G == 64, "groups of 8, 16, 32 or 64 lanes");
  PPC_DI static int lane() { return (int)__lane_id() & (G - 1); }
  // all-reduce sum / max over the group, bitwise identical in every lane
  template <typename T>
  PPC_DI static T sum(T v) {
    v += wave::dpp<wave::DPP_XOR1>(v);
    v += wave::dpp<wave::DPP_XOR2>(v);
    v += wave::dpp<wave::DPP_HALF_MIRROR>(v);
    if (G >= 16) v += wave::dpp<wave::DPP_MIRROR>(v);
    if (G >= 32) v += __shfl_xor(v, 16, 64);
    if (G >= 64) v += __shfl_xor(v, 32, 64);
    return v;
  }
  template <typename T>
  PPC_DI static T max(T v) {
    v = fmax(v, wave::dpp<wave::DPP_XOR1>(v));
    v = fmax(v, wave::dpp<wave::DPP_XOR2>(v));
    v = fmax(v, wave::dpp<wave::DPP_HALF_MIRROR>(v));
    if (G >= 16) v = fmax(v, wave::dpp<wave::DPP_MIRROR>(v));
    if (G >= 32) v = fmax(v, __shfl_xor(v, 16, 64));
    if (G >= 64) v = fmax(v, __shfl_xor(v, 32, 64));
    return v;
  }
  PPC_DI static unsigned long long umax(unsigned long long v) {
#pragma unroll
    for (int m = 1; m < G; m <<= 1) {
      const unsigned long long o = __shfl_xor(v, m, G);
      v = o > v ? o : v;
    }
    return v;
  }
  // value of lane src (0 <= src < G) of the group
  template <typename T>
  PPC_DI static T bcast(T v, int src) {
    return __shfl(v, src, G);
  }
};

// Orders this lane's LDS accesses around an exchange with the other lanes of
// its group (all in the same wavefront).
PPC_DI void group_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Loads and stores of the per-group Jacobian accumulator in the workspace,
// which the group's lanes update with device-scope atomics.
PPC_DI double ws_load(const double* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
PPC_DI void ws_store(double* p, double v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// ------------------------------------------- the mechanism on a group --
template <class T, int G>
struct GroupChem {
  using TM = TabMech<T>;
  using Gr = Grp<G>;
  static constexpr int NS = T::NS, NR = T::NR, n = T::NEQ;
  static constexpr int EPL = (n + G - 1) / G;   // equation slots per lane
  static constexpr int SPL = (NS + G - 1) / G;  // species slots per lane
  static constexpr int RPL = (NR + G - 1) / G;  // reaction slots per lane
  static constexpr int TL = NS % G, TS = NS / G;  // lane and slot of the temperature equation

  // per-group LDS: RHS exchange, Jacobian species data, the pivot row of
  // the Newton-matrix inversion and two vectors for its column scaling and
  // solves
  struct Shared {
    double c[NS], a[NS], b[NS], q[NR];
    double g[NS], h[NS], wd[NS];
    double v[n], w[n], prow[n];
  };

  // Per-lane constants of the lane's species, set once per kernel: when a
  // lane holds one species (SPL == 1) its NASA7 coefficients and range
  // midpoint (no table loads in the hot loop); for every slot, imw_k times
  // h_k/RT and cp_k/R (or e_k/RT, cv_k/R) at the bounds 90 K and 4000 K of
  // the temperature inversion.
  static constexpr int NC = SPL == 1 ? 1 : 0;
  struct Cache {
    double lo[NC > 0 ? NC : 1][7], hi[NC > 0 ? NC : 1][7], tmid[NC > 0 ? NC : 1];
    double hb[2][SPL], cb[2][SPL];
  };
  PPC_DI static void init_cache(Cache& cc, const int g, const bool enth) {
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      const int k = (e * G + g < NS) ? e * G + g : 0;
#pragma unroll
      for (int i = 0; i < 7; ++i) {
        cc.lo[e][i] = T::nasaLo_(7 * k + i);
        cc.hi[e][i] = T::nasaHi_(7 * k + i);
      }
      cc.tmid[e] = T::tmid_(k);
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const double Tb = b == 0 ? 90.0 : 4000.0;
#pragma unroll
      for (int e = 0; e < SPL; ++e) {
        const int k = e * G + g;
        cc.hb[b][e] = cc.cb[b][e] = 0.0;
        if (k < NS) {
          double a[7];
          TM::nasa(k, Tb, a);
          const double hp = TM::h_poly(a, Tb, 1.0 / Tb), cp = TM::cp_poly(a, Tb);
          cc.hb[b][e] = T::imw_(k) * (enth ? hp : hp - 1.0);
          cc.cb[b][e] = T::imw_(k) * (enth ? cp : cp - 1.0);
        }
      }
    }
  }
  // NASA7 coefficients of slot e (species k) for temperature Tv
  PPC_DI static void nasa(const Cache& cc, const int e, const int k, const double Tv, double* a) {
    if constexpr (NC > 0) {
      const bool lo = Tv < cc.tmid[0];
#pragma unroll
      for (int i = 0; i < 7; ++i) a[i] = lo ? cc.lo[0][i] : cc.hi[0][i];
    } else {
      TM::nasa(k, Tv, a);
    }
  }

  // mixture h (erg/g) and cp (erg/g/K), or e and cv, at Tv (Fuego.H CKHBMS/CKCPBS)
  template <bool ENTH>
  PPC_DI static void mix(const Cache& cc, const int g, const double (&Y)[SPL], const double Tv, double& hm,
                         double& cm) {
    const double invT = 1.0 / Tv;
    double hs = 0.0, cs = 0.0;
#pragma unroll
    for (int e = 0; e < SPL; ++e) {
      const int k = e * G + g;
      if (k < NS) {
        double a[7];
        nasa(cc, e, k, Tv, a);
        double hp = TM::h_poly(a, Tv, invT), cp = TM::cp_poly(a, Tv);
        if (!ENTH) {
          hp -= 1.0;
          cp -= 1.0;
        }
        const double yw = Y[e] * T::imw_(k);
        hs += yw * hp;
        cs += yw * cp;
      }
    }
    hm = Gr::sum(hs) * (RU * Tv);
    cm = Gr::sum(cs) * RU;
  }

  // mixture h (or e) at bound b (0: 90 K, 1: 4000 K) from the cached species
  // values; with cp too only when the caller extrapolates
  PPC_DI static double mix_bound(const Cache& cc, const int b, const double (&Y)[SPL], const bool want_c,
                                 double& cm) {
    double hs = 0.0, cs = 0.0;
#pragma unroll
    for (int e = 0; e < SPL; ++e) {
      hs += Y[e] * cc.hb[b][e];
      cs += Y[e] * cc.cb[b][e];
    }
    cm = want_c ? Gr::sum(cs) * RU : 0.0;
    return Gr::sum(hs) * (RU * (b == 0 ? 90.0 : 4000.0));
  }

  // GET_T_GIVEN_HY / GET_T_GIVEN_EY (Mechanisms/drm19/mechanism.H:2789-2890)
  // with group sums: Newton on T, |dT| clamped to 100 K, stop at |dT| < 1e-6,
  // at most 200 iterations, linear extrapolation outside [90, 4000] K.  (The
  // cache was built for the same energy form.)
  template <bool ENTH>
  PPC_DI static double T_from_energy(const Cache& cc, const int g, const double hv, const double (&Y)[SPL],
                                     const double tguess) {
    const double tmin = 90.0, tmax = 4000.0;
    double cmin, cmax;
    const double hmin = mix_bound(cc, 0, Y, false, cmin), hmax = mix_bound(cc, 1, Y, false, cmax);
    if (hv < hmin) {
      mix_bound(cc, 0, Y, true, cmin);
      return tmin - (hmin - hv) / cmin;
    }
    if (hv > hmax) {
      mix_bound(cc, 1, Y, true, cmax);
      return tmax - (hmax - hv) / cmax;
    }
    double t1 = tguess;
    if (t1 < tmin || t1 > tmax) t1 = tmin + (tmax - tmin) / (hmax - hmin) * (hv - hmin);
    for (int i = 0; i < 200; ++i) {
      double h1, c1;
      mix<ENTH>(cc, g, Y, t1, h1, c1);
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

  // net rates of progress of this lane's reactions into sh.q, from sh.c/a/b
  PPC_DI static void progress(Shared& sh, const int g, const double mixture, const double Tv, const double invT,
                              const double logT) {
    const double refC = 101325.0 / 8.31446 * invT, refCinv = 1.0 / refC;
#pragma unroll 1
    for (int r = 0; r < RPL; ++r) {
      const int j = r * G + g;
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
  }

  // fKernelSpec (Source/Reactions/ReactorUtils.H:358-415) for the group's
  // cell: y (this lane's equation slots), f = ydot of the same slots.
  PPC_DI static void rhs(Shared& sh, const Cache& cc, const int g, const int reactor_type, const double dt_save,
                         const double (&y)[EPL], const double (&src)[SPL], const double rhoe_init,
                         const double rhoesrc_ext, double (&f)[EPL]) {
    double rsum = 0.0;
#pragma unroll
    for (int e = 0; e < SPL; ++e)
      if (e * G + g < NS) rsum += y[e];
    const double rho = Gr::sum(rsum);
    const double rinv = 1.0 / rho;
    double Y[SPL];
#pragma unroll
    for (int e = 0; e < SPL; ++e) Y[e] = (e * G + g < NS) ? y[e] * rinv : 0.0;
    const double nrg = (rhoe_init + rhoesrc_ext * dt_save) * rinv;
    double Tv = Gr::bcast(y[TS], TL);
    const bool eform = reactor_type == E_REACTOR;
    Tv = eform ? T_from_energy<false>(cc, g, nrg, Y, Tv) : T_from_energy<true>(cc, g, nrg, Y, Tv);
    const double invT = 1.0 / Tv, logT = log(Tv), RT = RU * Tv;
    double ek[SPL], csum = 0.0, msum = 0.0;
#pragma unroll
    for (int e = 0; e < SPL; ++e) {
      const int k = e * G + g;
      ek[e] = 0.0;
      if (k < NS) {
        double a[7];
        nasa(cc, e, k, Tv, a);
        const double hp = TM::h_poly(a, Tv, invT), cp = TM::cp_poly(a, Tv);
        const double imw = T::imw_(k);
        ek[e] = (eform ? hp - 1.0 : hp) * RT * imw;  // CKUMS / CKHMS
        csum += Y[e] * imw * (eform ? cp - 1.0 : cp);
        const double sc = (rho * Y[e] * imw) * 1.0e6;  // CKYTCR, SI
        const double eg = exp(TM::g_poly(a, Tv, invT, logT));
        sh.c[k] = sc;
        sh.a[k] = 1.0 / eg;
        sh.b[k] = sc * eg;
        msum += sc;
      }
    }
    const double cmix = Gr::sum(csum) * RU;
    const double mixture = Gr::sum(msum);
    group_sync();
    progress(sh, g, mixture, Tv, invT, logT);
    group_sync();
    double esum = 0.0;
#pragma unroll
    for (int e = 0; e < EPL; ++e) f[e] = 0.0;
#pragma unroll
    for (int e = 0; e < SPL; ++e) {
      const int k = e * G + g;
      if (k < NS) {
        double w = 0.0;
        const int e1 = T::spPtr_(k + 1);
#pragma unroll 1
        for (int s = T::spPtr_(k); s < e1; ++s) w += T::spNu_(s) * sh.q[T::spRx_(s)];
        const double v = w * 1.0e-6 * T::mw_(k) + src[e];  // RTY2WDOT (CKWC units) + source
        f[e] = v;
        esum += v * ek[e];
      }
    }
    group_sync();
    const double srcT = rhoesrc_ext - Gr::sum(esum);
    if (g == TL) f[TS] = srcT * (rinv / cmix);
  }

  // The BDF Jacobian d f_i / d y_j at y (bdf_jacobian in chem_cell.h with
  // s = 1): accumulated in the group's n x n FP64 area Jw of the workspace
  // (column-major, aJacobian / DWDOT layout), then row i = e*G + g written
  // to Jrow(e, j) in JT.
  template <int CONSP, typename JT, class JR>
  PPC_DI static void jacobian(Shared& sh, const Cache& cc, const int g, double* Jw, const double (&y)[EPL],
                              JR Jrow) {
    const double Tv = Gr::bcast(y[TS], TL);
    const double invT = 1.0 / Tv, invT2 = invT * invT, logT = log(Tv);
    const double refC = 101325.0 / 8.31446 * invT, refCinv = 1.0 / refC;
    double cR[SPL], dcR[SPL], eh[SPL], sc[SPL];
    double msum = 0.0;
#pragma unroll
    for (int e = 0; e < SPL; ++e) {
      const int k = e * G + g;
      cR[e] = dcR[e] = eh[e] = sc[e] = 0.0;
      if (k < NS) {
        double a[7];
        nasa(cc, e, k, Tv, a);
        const double hRT = TM::h_poly(a, Tv, invT), cpR = TM::cp_poly(a, Tv);
        sc[e] = y[e] * T::imw_(k) * 1.0e6;  // CKYTCR + DWDOT (x1e6)
        sh.c[k] = sc[e];
        sh.g[k] = TM::g_poly(a, Tv, invT, logT);
        sh.h[k] = hRT;
        sh.wd[k] = 0.0;
        cR[e] = (CONSP == 1) ? cpR : cpR - 1.0;
        eh[e] = (CONSP == 1) ? hRT : hRT - 1.0;
        dcR[e] = TM::dcp_poly(a, Tv);
        msum += sc[e];
      }
    }
#pragma unroll 1
    for (int i = g; i < n * n; i += G) ws_store(Jw + i, 0.0);
    const double mixture = Gr::sum(msum);
    __threadfence();
    group_sync();
#pragma unroll 1
    for (int r = 0; r < RPL; ++r) {
      const int j = r * G + g;
      if (j < NR)
        TM::template jac_reaction<CONSP>(
            j, sh.c, sh.g, sh.h, mixture, Tv, invT, invT2, logT, refC, refCinv,
            [&](int idx, double v) { unsafeAtomicAdd(Jw + idx, v); },
            [&](int m, double v) { atomicAdd(&sh.wd[m], v); });
    }
    __threadfence();
    group_sync();
    // temperature row (aJacobian, Mechanisms/drm19/mechanism.H:15317-15352)
    double s_cm = 0.0, s_dcm = 0.0, s_eh = 0.0, s_deh = 0.0;
#pragma unroll
    for (int e = 0; e < SPL; ++e) {
      const int k = e * G + g;
      if (k < NS) {
        const double wdk = sh.wd[k];
        s_cm += cR[e] * sc[e];
        s_dcm += dcR[e] * sc[e];
        s_eh += eh[e] * wdk;
        s_deh += invT * (cR[e] - eh[e]) * wdk + eh[e] * ws_load(Jw + NS * n + k);
      }
    }
    const double cmix = Gr::sum(s_cm), dcmixdT = Gr::sum(s_dcm);
    const double ehmix = Gr::sum(s_eh), dehmixdT = Gr::sum(s_deh);
    const double cmixinv = 1.0 / cmix;
    const double tmp1 = ehmix * cmixinv, tmp3 = cmixinv * Tv, tmp2 = tmp1 * tmp3;
    double trow[SPL];
#pragma unroll
    for (int e = 0; e < SPL; ++e) {
      const int k = e * G + g;
      trow[e] = 0.0;
      if (k < NS) {
        double d = 0.0;
#pragma unroll 1
        for (int mm = 0; mm < NS; ++mm) d += ((CONSP == 1) ? sh.h[mm] : sh.h[mm] - 1.0) * ws_load(Jw + k * n + mm);
        trow[e] = tmp2 * cR[e] - tmp3 * d;
      }
    }
    __threadfence();
    group_sync();
#pragma unroll
    for (int e = 0; e < SPL; ++e)
      if (e * G + g < NS) ws_store(Jw + (e * G + g) * n + NS, trow[e]);
    if (g == 0) ws_store(Jw + n * n - 1, -tmp1 + tmp2 * dcmixdT - tmp3 * dehmixdT);
    __threadfence();
    group_sync();
    // DWDOT scaling (T column x1e-6, T row x1e6) and the transpose with
    // molecular-weight scaling: row i, column j = d f_i / d y_j
#pragma unroll
    for (int e = 0; e < EPL; ++e) {
      const int i = e * G + g;
      if (i < n) {
        const bool spi = i < NS;
        const double wi = spi ? T::mw_(i) : 1.0;
#pragma unroll 1
        for (int jj = 0; jj < n; ++jj) {
          double a;
          if (jj < NS) {
            const double Jv = spi ? ws_load(Jw + jj * n + i) : ws_load(Jw + jj * n + NS) * 1.0e6;
            a = Jv * wi * T::imw_(jj);
          } else {
            const double Jv = spi ? ws_load(Jw + NS * n + i) * 1.0e-6 : ws_load(Jw + n * n - 1);
            a = Jv * wi;
          }
          Jrow(e, jj) = (JT)a;
        }
      }
    }
  }
};

// ------------------------------------------- group Newton-matrix inverse --
// Gauss-Jordan exchange steps on the row-distributed n x n matrix a (lane g
// holds rows i = e*G + g, e < EPL) with partial pivoting: step k pivots on the
// largest |a(i, k)| among the rows not yet used (the choice LU with partial
// pivoting makes), lowest row on ties; rows never move.  The pivot row goes
// through the group's LDS (prow).  Exchange on (r, k): a(r, k) <- 1 / a(r, k),
// a(r, j) <- -a(r, j) / a(r, k), a(i, k) <- a(i, k) / a(r, k) and
// a(i, j) <- a(i, j) - a(i, k) a(r, j) / a(r, k) for the other rows.  After the
// n steps (kstep[e] = the step at which row i was the pivot) A x = b is
// solved by x_{kstep(i)} = sum_j a(i, j) b_{p_j}, p_j the pivot row of step j:
// one dot product per row instead of two sequential triangular sweeps.
template <int n, int G, int EPL, typename LT>
PPC_DI void group_inv_factor(LT (&a)[EPL][n], int (&kstep)[EPL], double* prow, const int g) {
  using Gr = Grp<G>;
#pragma unroll
  for (int e = 0; e < EPL; ++e) kstep[e] = (e * G + g < n) ? -1 : n;
  group_sync();
#pragma unroll
  for (int k = 0; k < n; ++k) {
    unsigned long long key = 0;
#pragma unroll
    for (int e = 0; e < EPL; ++e) {
      if (kstep[e] < 0) {
        const float m = (float)fabs(a[e][k]);
        const unsigned long long kk =
            ((unsigned long long)__float_as_uint(m) << 32) | (unsigned long long)(0xffffffffu - (unsigned)(e * G + g));
        key = kk > key ? kk : key;
      }
    }
    key = Gr::umax(key);
    const int p = (int)(0xffffffffu - (unsigned)(key & 0xffffffffu));
    const int pl = p % G, ps = p / G;
    if (g == pl) {
#pragma unroll
      for (int j = 0; j < n; ++j) {
        LT v = a[0][j];
#pragma unroll
        for (int e = 1; e < EPL; ++e) v = (ps == e) ? a[e][j] : v;
        prow[j] = (double)v;
      }
    }
    group_sync();
    const LT prs = LT(1) / (LT)prow[k];
#pragma unroll
    for (int e = 0; e < EPL; ++e) {
      const bool isp = e * G + g == p;
      kstep[e] = isp ? k : kstep[e];
      const LT f = a[e][k] * prs;
#pragma unroll
      for (int j = 0; j < n; ++j) {
        if (j == k) continue;
        const LT r = (LT)prow[j];
        a[e][j] = isp ? -r * prs : a[e][j] - f * r;
      }
      a[e][k] = isp ? prs : f;
    }
    group_sync();
  }
}

// x (this lane's equation slots) = A^-1 b (this lane's rows) from the
// exchanged matrix of group_inv_factor; v and w are n-vectors of the group's
// LDS.
template <int n, int G, int EPL, typename LT>
PPC_DI void group_inv_solve(const LT (&a)[EPL][n], const int (&kstep)[EPL], double* v, double* w, const int g,
                            const double (&b)[EPL], double (&x)[EPL]) {
#pragma unroll
  for (int e = 0; e < EPL; ++e)
    if (e * G + g < n) v[kstep[e]] = b[e];  // v_j = b_{p_j}
  group_sync();
  double acc[EPL];
#pragma unroll
  for (int e = 0; e < EPL; ++e) acc[e] = 0.0;
#pragma unroll
  for (int j = 0; j < n; ++j) {
    const double vj = v[j];
#pragma unroll
    for (int e = 0; e < EPL; ++e) acc[e] += (double)a[e][j] * vj;
  }
#pragma unroll
  for (int e = 0; e < EPL; ++e)
    if (e * G + g < n) w[kstep[e]] = acc[e];  // x_{kstep(i)}
  group_sync();
#pragma unroll
  for (int e = 0; e < EPL; ++e) x[e] = (e * G + g < n) ? w[e * G + g] : 0.0;
  group_sync();
}

// D[0..order][e] <- (R U)^T D[0..order][e] for a step-size ratio `factor`
// (change_D in reactors.h) on this lane's equation slots, as U^T (R^T D):
// R[i][j] = prod_{m=1..i} (m - 1 - factor j) / m (R[i][0] = [i == 0]) and
// U = R at factor 1, both formed on the fly so that only the vector R^T D is
// held.
template <int EPL>
PPC_DI void group_change_D(double (&D)[AMAXORD + 3][EPL], const int order, const double factor) {
  double w[AMAXORD + 1][EPL];
#pragma unroll
  for (int k = 0; k <= AMAXORD; ++k) {
    double r = 1.0;  // R[kp][k]
#pragma unroll
    for (int e = 0; e < EPL; ++e) w[k][e] = D[0][e];
#pragma unroll
    for (int kp = 1; kp <= AMAXORD; ++kp) {
      r *= (k == 0) ? 0.0 : (kp - 1 - factor * k) / (double)kp;
#pragma unroll
      for (int e = 0; e < EPL; ++e)
        if (kp <= order) w[k][e] += r * D[kp][e];
    }
  }
#pragma unroll
  for (int i = 0; i <= AMAXORD; ++i) {
    double u = 1.0;  // U[k][i]
    double s[EPL];
#pragma unroll
    for (int e = 0; e < EPL; ++e) s[e] = w[0][e];
#pragma unroll
    for (int k = 1; k <= AMAXORD; ++k) {
      u *= (i == 0) ? 0.0 : (k - 1 - i) / (double)k;
#pragma unroll
      for (int e = 0; e < EPL; ++e)
        if (k <= order) s[e] += u * w[k][e];
    }
#pragma unroll
    for (int e = 0; e < EPL; ++e)
      if (i <= order) D[i][e] = s[e];
  }
}

// Doubles of device workspace per group: the FP64 Jacobian accumulator and
// the group's Jacobian rows in JT.
template <int n, int G, typename JT>
constexpr long long group_ws_doubles() {
  constexpr long long epl = (n + G - 1) / G;
  return (long long)n * n + (epl * n * G * (long long)sizeof(JT) + 7) / 8;
}

// Persistent kernel: each group of G lanes integrates cells from the queue
// until it is empty.  ws: (gridDim.x * 64 / G) group areas of
// group_ws_doubles<NEQ, G, LU_T>() doubles.
template <class T, typename LU_T, int G, int OCC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
group_bdf_kernel(ReactParamsDev P, IO1D io, double dt, QueueArgs q, CellStats st) {
  using C = GroupChem<T, G>;
  using Gr = Grp<G>;
  constexpr int n = C::n, NS = C::NS, EPL = C::EPL, SPL = C::SPL, NG = 64 / G;
  __shared__ typename C::Shared shg[NG];
  const int g = Gr::lane();
  const int gi = (int)__lane_id() / G;
  typename C::Shared& sh = shg[gi];
  typename C::Cache cc;
  C::init_cache(cc, g, P.reactor_type != E_REACTOR);
  const long long gid = (long long)blockIdx.x * NG + gi;
  double* Jw = P.ws + gid * group_ws_doubles<n, G, LU_T>();
  // this lane's Jacobian rows: element (e, j) at jrow[(e * n + j) * G]
  LU_T* const jrow = reinterpret_cast<LU_T*>(Jw + (long long)n * n) + g;
  auto Jrow = [&](int e, int j) -> LU_T& { return jrow[(e * n + j) * G]; };

  const AdaptParamsDev& ap = P.adapt;
  const double rtol = ap.rtol, atol = ap.atol;
  const double newton_tol = fmax(10 * 2.220446049250313e-16 / rtol, fmin(0.03, sqrt(rtol)));
  const double t_bound = dt;
  const long long rhs_cap = rhs_budget(ap.max_steps);
  const int lu_batch = ap.lu_batch > 0 ? (ap.lu_batch < NG ? ap.lu_batch : NG) : NG;
  const int lu_wait = ap.lu_wait > 0 ? ap.lu_wait : 32;

  // this lane's equation slots (rows i = e*G + g); `valid` marks real equations
  bool valid[EPL];
#pragma unroll
  for (int e = 0; e < EPL; ++e) valid[e] = e * G + g < n;
  double D[AMAXORD + 3][EPL];
  double yv[EPL], f[EPL], d[EPL], ypred[EPL], psi[EPL], scl[EPL], lus[EPL], yjac[EPL], src[SPL];
  LU_T LU[EPL][n];  // the exchanged (inverted) Newton matrix, this lane's rows
  int kstep[EPL];
  double rhoe = 0, rhoesrc = 0;
  double t = 0, t_new = 0, h_abs = 0, c = 0, dy_norm_old = -1, h0 = 0;
  int order = 1, n_equal = 0, k_newton = 0, phase = PH_F0;
  bool lu_valid = false, current_jac = false, need_jac = false, need_lu = false;
  int nsteps = 0, nrhs = 0, njac = 0, nlu = 0, nnewton = 0, nrej = 0, status = 0, wait = 0;
  int cell = 0, cid = 0;
  bool active = false;

  // root-mean-square norm over the cell's n equations
  auto rms = [&](const double (&x)[EPL], const double (&s)[EPL]) {
    double acc = 0.0;
#pragma unroll
    for (int e = 0; e < EPL; ++e)
      if (valid[e]) {
        const double v = x[e] / s[e];
        acc += v * v;
      }
    return sqrt(Gr::sum(acc) / n);
  };

  auto next_cell = [&]() {
    int cl = 0;
    if (g == 0) cl = atomicAdd(q.counter, 1);
    cell = Gr::bcast(cl, 0);
    active = cell < q.ncells;
    if (!active) return;
    cid = q.order ? q.order[cell] : cell;
#pragma unroll
    for (int e = 0; e < EPL; ++e) {
      const int i = e * G + g;
      D[0][e] = valid[e] ? io.y[(long long)cid * n + i] : 0.0;
      yv[e] = D[0][e];
#pragma unroll
      for (int k = 1; k < AMAXORD + 3; ++k) D[k][e] = 0.0;
      f[e] = d[e] = ypred[e] = psi[e] = 0.0;
      scl[e] = lus[e] = 1.0;
    }
#pragma unroll
    for (int e = 0; e < SPL; ++e) {
      const int k = e * G + g;
      src[e] = k < NS ? io.ysrc[(long long)cid * NS + k] : 0.0;
    }
    rhoe = io.rX[cid];
    rhoesrc = io.rXsrc[cid];
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
#pragma unroll
    for (int e = 0; e < EPL; ++e)
      if (valid[e]) io.y[(long long)cid * n + e * G + g] = D[0][e];
    if (g == 0) {
      io.rX[cid] = rhoe + dt * rhoesrc;
      st.nsteps[cid] = status == 0 ? nsteps : status;
      if (st.nrhs) {
        st.nrhs[cid] = nrhs;
        st.njac[cid] = njac;
        st.nlu[cid] = nlu;
        st.nnewton[cid] = nnewton;
        st.nreject[cid] = nrej;
      }
    }
    next_cell();
  };

  next_cell();

  // S^-1 (I - c J) S in LU_T with S the current error weights, factored
  auto factor = [&]() {
#pragma unroll
    for (int e = 0; e < EPL; ++e) {
      lus[e] = scl[e];
      if (valid[e]) sh.v[e * G + g] = lus[e];
    }
    group_sync();
#pragma unroll
    for (int e = 0; e < EPL; ++e) {
      const double si = 1.0 / lus[e];
      const int i = e * G + g;
#pragma unroll
      for (int j = 0; j < n; ++j) {
        const double jf = valid[e] ? (double)Jrow(e, j) : 0.0;
        const double m = (i == j ? 1.0 : 0.0) - c * jf;
        LU[e][j] = valid[e] ? (LU_T)(si * m * sh.v[j]) : LU_T(0);
      }
    }
    group_inv_factor<n, G, EPL, LU_T>(LU, kstep, sh.prow, g);
    nlu++;
    lu_valid = true;
  };

  auto jacobian = [&](const double (&yj)[EPL]) {
    if (P.reactor_type == H_REACTOR)
      C::template jacobian<1, LU_T>(sh, cc, g, Jw, yj, Jrow);
    else
      C::template jacobian<0, LU_T>(sh, cc, g, Jw, yj, Jrow);
    __threadfence();
    group_sync();
    njac++;
  };

  auto request_jacobian = [&](const double (&yj)[EPL]) {
#pragma unroll
    for (int e = 0; e < EPL; ++e) yjac[e] = yj[e];
    need_jac = true;
  };

  auto begin_attempt = [&]() -> bool {
    const double min_step = 10 * fabs(nextafter(t, INFINITY) - t);
    if (!(h_abs >= min_step)) {  // also ends a cell whose step size is NaN
      status = -3;
      return false;
    }
    t_new = t + h_abs;
    if (t_new - t_bound > 0) {
      t_new = t_bound;
      group_change_D<EPL>(D, order, fabs(t_new - t) / h_abs);
      n_equal = 0;
      lu_valid = false;
    }
    const double h = t_new - t;
    h_abs = fabs(h);
    const double inv_alpha = 1.0 / ndf_alpha(order);
#pragma unroll
    for (int e = 0; e < EPL; ++e) {
      double sum = 0.0, ps = 0.0;
#pragma unroll
      for (int kk = 0; kk <= AMAXORD; ++kk)
        if (kk <= order) sum += D[kk][e];
#pragma unroll
      for (int kk = 1; kk <= AMAXORD; ++kk)
        if (kk <= order) ps += D[kk][e] * bdf_gamma(kk);
      ypred[e] = sum;
      scl[e] = atol + rtol * fabs(sum);
      psi[e] = ps * inv_alpha;
      yv[e] = sum;
      d[e] = 0.0;
    }
    c = h * inv_alpha;
    if (!lu_valid) need_lu = true;
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
      group_change_D<EPL>(D, order, min_step / h_abs);
      h_abs = min_step;
      n_equal = 0;
    }
    current_jac = false;
    return begin_attempt();
  };

  while (active) {
    // ---- Jacobian / LU requests, served for the waiting groups together
    const unsigned long long lead = __ballot(g == 0);
    const unsigned long long waiting = __ballot(g == 0 && need_lu);
    if (waiting) {
      const bool stale = __ballot(g == 0 && need_lu && wait >= lu_wait) != 0;
      if (__popcll(waiting) >= lu_batch || waiting == lead || stale) {
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
    C::rhs(sh, cc, g, P.reactor_type, t_eval, yv, src, rhoe, rhoesrc, f);
    nrhs++;
    bool ok = true;
    if (phase == PH_F0) {
      // f0 = f(y0): initial step size (Hairer, Norsett & Wanner I, II.4)
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        D[1][e] = f[e];  // f0 kept in D[1] until h is known
        scl[e] = atol + fabs(D[0][e]) * rtol;
      }
      if (ap.first_step > 0) {
        h_abs = ap.first_step;
#pragma unroll
        for (int e = 0; e < EPL; ++e) {
          D[1][e] *= h_abs;
          yv[e] = D[0][e];
        }
        request_jacobian(yv);
        ok = begin_step();
      } else {
        const double d0 = rms(D[0], scl), d1 = rms(f, scl);
        h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
        h0 = fmin(h0, t_bound);
#pragma unroll
        for (int e = 0; e < EPL; ++e) yv[e] = D[0][e] + h0 * f[e];
        phase = PH_F1;
      }
    } else if (phase == PH_F1) {
#pragma unroll
      for (int e = 0; e < EPL; ++e) d[e] = f[e] - D[1][e];
      const double d1 = rms(D[1], scl);
      const double d2 = rms(d, scl) / h0;
      const double h1 = (d1 <= 1e-15 && d2 <= 1e-15) ? fmax(1e-6, h0 * 1e-3) : sqrt(0.01 / fmax(d1, d2));
      h_abs = fmin(fmin(100 * h0, h1), t_bound);
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        D[1][e] *= h_abs;
        yv[e] = D[0][e];
      }
      request_jacobian(yv);
      ok = begin_step();
    } else {
      // one modified-Newton iteration with f = f(t_new, yv)
      nnewton++;
      double bad = 0.0;
#pragma unroll
      for (int e = 0; e < EPL; ++e)
        if (valid[e] && !isfinite(f[e])) bad = 1.0;
      const bool finite = Gr::max(bad) == 0.0;
      bool failed = !finite, converged = false;
      const int n_iter = k_newton + 1;
      if (finite) {
        double b[EPL], dy[EPL];
#pragma unroll
        for (int e = 0; e < EPL; ++e) b[e] = valid[e] ? (c * f[e] - psi[e] - d[e]) / lus[e] : 0.0;
        group_inv_solve<n, G, EPL, LU_T>(LU, kstep, sh.v, sh.w, g, b, dy);
#pragma unroll
        for (int e = 0; e < EPL; ++e) dy[e] *= lus[e];
        const double dy_norm = rms(dy, scl);
        const double rate = dy_norm_old < 0 ? -1.0 : dy_norm / dy_norm_old;
        if (rate >= 0 && (rate >= 1 || pow(rate, 4 - k_newton) / (1 - rate) * dy_norm > newton_tol)) {
          failed = true;
        } else {
#pragma unroll
          for (int e = 0; e < EPL; ++e) {
            yv[e] += dy[e];
            d[e] += dy[e];
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
#pragma unroll
          for (int e = 0; e < EPL; ++e) {
            yv[e] = ypred[e];
            d[e] = 0.0;
          }
          dy_norm_old = -1.0;
          k_newton = 0;
        } else {
          h_abs *= 0.5;
          group_change_D<EPL>(D, order, 0.5);
          n_equal = 0;
          lu_valid = false;
          nrej++;
          ok = begin_attempt();
        }
      } else if (converged) {
        const double safety = 0.9 * 9.0 / (double)(8 + n_iter);
        double err[EPL];
#pragma unroll
        for (int e = 0; e < EPL; ++e) {
          scl[e] = atol + rtol * fabs(yv[e]);
          err[e] = ndf_econst(order) * d[e];
        }
        const double error_norm = rms(err, scl);
        if (error_norm > 1) {
          const double fac = fmax(0.2, safety * pow(error_norm, -1.0 / (order + 1)));
          h_abs *= fac;
          group_change_D<EPL>(D, order, fac);
          n_equal = 0;
          nrej++;
          ok = begin_attempt();
        } else {
          // accept: D[order+2] = d - D[order+1]; D[order+1] = d; D[k] += D[k+1], k = order..0
          nsteps++;
          n_equal++;
          t = t_new;
#pragma unroll
          for (int e = 0; e < EPL; ++e) {
#pragma unroll
            for (int kk = AMAXORD + 2; kk >= 1; --kk) {
              if (kk == order + 2) D[kk][e] = d[e] - D[kk - 1][e];
              if (kk == order + 1) D[kk][e] = d[e];
            }
#pragma unroll
            for (int kk = AMAXORD; kk >= 0; --kk)
              if (kk <= order) D[kk][e] += D[kk + 1][e];
          }
          if (n_equal >= order + 1) {
            double em = INFINITY, ep = INFINITY;
            if (order > 1) {
#pragma unroll
              for (int e = 0; e < EPL; ++e) {
                double v = 0.0;
#pragma unroll
                for (int kk = 0; kk <= AMAXORD; ++kk)
                  if (kk == order) v = D[kk][e];
                err[e] = ndf_econst(order - 1) * v;
              }
              em = rms(err, scl);
            }
            if (order < AMAXORD) {
#pragma unroll
              for (int e = 0; e < EPL; ++e) {
                double v = 0.0;
#pragma unroll
                for (int kk = 0; kk < AMAXORD + 3; ++kk)
                  if (kk == order + 2) v = D[kk][e];
                err[e] = ndf_econst(order + 1) * v;
              }
              ep = rms(err, scl);
            }
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
            group_change_D<EPL>(D, order, fac);
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
