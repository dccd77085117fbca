// Box (Array4) <-> flat (YCOrder) cell state, the pre- and post-processing of
// the reference's 1-D callers (utils::box_flatten / box_unflatten,
// Source/Reactions/ReactorUtils.H:160-289, FlattenOps :291-356), one thread
// per cell:
//
//  flatten:   y[c*(N+1)+n] = rhoY_n (clipped to [0, rho_old] with
//             clean_init_massfrac), ysrc[c*N+n] = rhoYsrc_n, e[c] = rhoE,
//             esrc[c] = rhoEsrc, y[c*(N+1)+N] = T from rhoE / rho with the
//             box T as the Newton guess;
//  unflatten: rhoY_n = y_n, or with clean_init_massfrac the box's old rhoY_n
//             plus the increment over the clipped value, scaled by
//             rho_old / rho_clipped; rhoE = e + dt * rhoEsrc; T from rhoE / rho
//             with the flat T as the guess; FC = fc[c].
//
// Cell c of the box [lo, lo + len) is c = (k - lo_z) len_x len_y + (j - lo_y) len_x + (i - lo_x).
// With a.divide the energy per unit mass is e / rho, as ReactorBDF's own box
// path forms it (ReactorBDF.cpp:379,469), instead of FlattenOps' e * (1 / rho):
// the batched box form runs through these kernels and the 1-D kernel.
#pragma once
#include "chem_cell.h"
#include "reactors_fwd.h"

namespace ppc {

template <class M>
__global__ void __launch_bounds__(256) flatten_kernel(FlatArgs a) {
  constexpr int NS = M::NS;
  const BoxArgs& b = a.b;
  const int ncells = b.len[0] * b.len[1] * b.len[2];
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
    const int i = b.lo[0] + c % b.len[0];
    const int j = b.lo[1] + (c / b.len[0]) % b.len[1];
    const int k = b.lo[2] + c / (b.len[0] * b.len[1]);
    double rho_old = 0.0;
    if (a.clean)
      for (int n = 0; n < NS; ++n) rho_old += b.rY(i, j, k, n);
    double rY[NS], Y[NS];
    for (int n = 0; n < NS; ++n) {
      const double v = b.rY(i, j, k, n);
      rY[n] = a.clean ? fmin(fmax(0.0, v), rho_old) : v;
      a.y[(long long)c * (NS + 1) + n] = rY[n];
      a.ysrc[(long long)c * NS + n] = b.rYsrc(i, j, k, n);
    }
    double rinv;
    const double rho = rho_and_Y<M>(rY, Y, rinv);
    const double e = b.rE(i, j, k, 0);
    a.e[c] = e;
    a.esrc[c] = b.rEsrc(i, j, k, 0);
    a.y[(long long)c * (NS + 1) + NS] =
        T_from_energy_rt<M>(a.reactor_type, a.divide ? e / rho : e * rinv, Y, b.T(i, j, k, 0));
  }
}

template <class M>
__global__ void __launch_bounds__(256) unflatten_kernel(FlatArgs a) {
  constexpr int NS = M::NS;
  const BoxArgs& b = a.b;
  const int ncells = b.len[0] * b.len[1] * b.len[2];
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
    const int i = b.lo[0] + c % b.len[0];
    const int j = b.lo[1] + (c / b.len[0]) % b.len[1];
    const int k = b.lo[2] + c / (b.len[0] * b.len[1]);
    const double* yc = a.y + (long long)c * (NS + 1);
    double rY[NS], Y[NS];
    if (a.clean) {
      double rho_old = 0.0, rho_clip = 0.0;
      for (int n = 0; n < NS; ++n) rho_old += b.rY(i, j, k, n);
      for (int n = 0; n < NS; ++n) rho_clip += fmin(fmax(0.0, b.rY(i, j, k, n)), rho_old);
      for (int n = 0; n < NS; ++n) {
        const double old = b.rY(i, j, k, n);
        rY[n] = old + (yc[n] - fmin(fmax(0.0, old), rho_old)) * rho_old / rho_clip;
      }
    } else {
      for (int n = 0; n < NS; ++n) rY[n] = yc[n];
    }
    for (int n = 0; n < NS; ++n) b.rY(i, j, k, n) = rY[n];
    double rinv;
    const double rho = rho_and_Y<M>(rY, Y, rinv);
    const double e = a.e[c] + a.dt * b.rEsrc(i, j, k, 0);
    b.rE(i, j, k, 0) = e;
    b.T(i, j, k, 0) = T_from_energy_rt<M>(a.reactor_type, a.divide ? e / rho : e * rinv, Y, yc[NS]);
    if (b.FC.p) b.FC(i, j, k, 0) = a.fc ? (double)a.fc[c] : (double)a.fc32[c];
  }
}

}  // namespace ppc
