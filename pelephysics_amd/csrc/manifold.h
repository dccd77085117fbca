// ReactorRK64 with the Manifold EOS: the state is a tabulated manifold of D
// parameters plus the density (Source/Eos/Manifold.H, Mechanisms/Null:
// NUM_SPECIES = D + 1, the last "species" is rho), every EOS query and the
// parameter source terms are multilinear interpolations in a table
// (Source/Utility/BlackBoxFunction/Table.H) or outputs of a small dense network
// (manifold.model = NeuralNet, NeuralNetHomerolled.H), evaluated per lane.
//
//  * lookup: the left grid index of each coordinate (bisection, clamped to
//    [0, len - 2]), alpha = (x_{i+1} - x) / (x_{i+1} - x_i) clamped to [0, 1],
//    value = alpha * left + (1 - alpha) * right, nested from the last
//    dimension down (Table.H:183-219, 470-545);
//  * RHS (fKernelSpec under USE_MANIFOLD_EOS, ReactorUtils.H:358-415,
//    Manifold.H:265-276): rho = y[D]; the D parameter rates are the tabulated
//    sources at Y = y / rho (zero where the table has none) plus the external
//    sources; the density rate is its external source; dT/dt = rhoE_src / rho
//    (species energies 0, Cv = 1: T is carried, not solved for);
//  * box form (ReactorRK64.cpp:166-320): T from the table at the initial and
//    at the final state (RHY2T / REY2T; 1 without compute_temperature).
#pragma once
#include "reactors.h"

namespace ppc {

// Left index and weight along dimension d at coordinate x.
PPC_DI void man_locate(const ManifoldDev& m, int d, double x, int& idx, double& alpha) {
  const double* g = m.grids + m.dim_off[d];
  const int n = m.dim_len[d];
  int lo = 0;
  if (x >= g[n - 2]) {
    lo = n - 2;
  } else if (x > g[0]) {
    int hi = n - 2;
    while (hi - lo > 1) {
      const int mid = (hi + lo) / 2;
      if (x >= g[mid]) {
        lo = mid;
      } else {
        hi = mid;
      }
    }
  }
  idx = lo;
  const double a = (g[lo + 1] - x) * (1.0 / (g[lo + 1] - g[lo]));
  alpha = fmin(fmax(a, 0.0), 1.0);
}

// Multilinear interpolation of the slab `data` over dimensions [0, L).
template <int L>
PPC_DI double man_interp(const ManifoldDev& m, const int* idx, const double* alpha, const double* data) {
  if constexpr (L == 1) {
    const double left = data[idx[0]], right = data[idx[0] + 1];
    return alpha[0] * left + (1.0 - alpha[0]) * right;
  } else {
    const int s = m.dim_stride[L - 1];
    const double left = man_interp<L - 1>(m, idx, alpha, data + (long long)idx[L - 1] * s);
    const double right = man_interp<L - 1>(m, idx, alpha, data + (long long)(idx[L - 1] + 1) * s);
    return alpha[L - 1] * left + (1.0 - alpha[L - 1]) * right;
  }
}

template <int D>
struct ManPoint {
  int idx[D];
  double alpha[D];
  PPC_DI ManPoint(const ManifoldDev& m, const double* Y) {
#pragma unroll
    for (int d = 0; d < D; ++d) man_locate(m, d, Y[d], idx[d], alpha[d]);
  }
  PPC_DI double value(const ManifoldDev& m, int ivar) const {
    return man_interp<D>(m, idx, alpha, m.values + (long long)ivar * m.var_stride);
  }
};

// NNModel::operator() (NeuralNetModelDef.H:227-249, the layers of
// NeuralNetLayerDef.H): Linear out_j = sum_i in_i W[j][i] (summed in i order)
// + b_j; LeakyReLU max-clamp plus slope x min-clamp (a NaN passes through);
// BatchNorm1d in_i scale_i + shift_i.  a and b hold MAN_NN_MAXW values each;
// returns the buffer with the outputs.
PPC_DI const double* man_nn_forward(const ManifoldDev& m, const double* x, double* a, double* b) {
  const double* in = x;
  double* out = a;
  for (int l = 0; l < m.nn_layers; ++l) {
    const double* p = m.nn + m.nn_off[l];
    const int nin = m.nn_nin[l], nout = m.nn_nout[l];
    if (m.nn_kind[l] == 0) {
      for (int j = 0; j < nout; ++j) {
        double s = 0.0;
        for (int i = 0; i < nin; ++i) s += in[i] * p[nin * j + i];
        out[j] = s + p[nin * nout + j];
      }
    } else if (m.nn_kind[l] == 1) {
      const double slope = p[0];
      for (int i = 0; i < nin; ++i) {
        const double v = in[i];
        out[i] = (v < 0.0 ? 0.0 : v) + slope * (v > 0.0 ? 0.0 : v);
      }
    } else {
      for (int i = 0; i < nin; ++i) out[i] = in[i] * p[i] + p[nin + i];
    }
    in = out;
    out = out == a ? b : a;
  }
  return in;
}

// BlackBoxFunction::get_values (BlackBoxFunction.H:63-67): the variables
// ivar[0..nreq) at the manifold point Y (a negative index gives 0), from the
// table or from one evaluation of the network.
template <int D>
PPC_DI void man_values(const ManifoldDev& m, const double* Y, int nreq, const int* ivar, double* out) {
  if (m.model == 1) {
    double a[MAN_NN_MAXW], b[MAN_NN_MAXW];
    const double* o = man_nn_forward(m, Y, a, b);
    for (int i = 0; i < nreq; ++i) out[i] = ivar[i] >= 0 ? o[ivar[i]] : 0.0;
  } else {
    const ManPoint<D> pt(m, Y);
    for (int i = 0; i < nreq; ++i) out[i] = ivar[i] >= 0 ? pt.value(m, ivar[i]) : 0.0;
  }
}

// Manifold::RHY2T / REY2T (Manifold.H:176-210).
template <int D>
PPC_DI double man_temperature(const ManifoldDev& m, const double* Y) {
  if (!m.compute_T) return 1.0;
  double T;
  man_values<D>(m, Y, 1, &m.idx_T, &T);
  return T;
}

// fKernelSpec with the Manifold EOS.
template <int D>
PPC_DI void man_rhs(const ManifoldDev& m, const double* __restrict__ y, double* __restrict__ ydot,
                    const double rhoesrc_ext, const double* __restrict__ rYsrc) {
  const double rho = y[D];
  const double rinv = 1.0 / rho;
  double Y[D];
#pragma unroll
  for (int d = 0; d < D; ++d) Y[d] = y[d] * rinv;
  double wdot[D];
  man_values<D>(m, Y, D, m.idx_src, wdot);
#pragma unroll
  for (int d = 0; d < D; ++d) ydot[d] = wdot[d] + rYsrc[d];
  ydot[D] = 0.0 + rYsrc[D];
  ydot[D + 1] = rhoesrc_ext * (rinv / 1.0);
}

template <int D>
PPC_DI int man_rk64_cell(const ReactParamsDev& P, double* __restrict__ y, const double rhoesrc,
                         const double* __restrict__ rYsrc, const double dt_react) {
  return rk64_core<D + 2>(P.rk64, y, dt_react,
                          [&](double, const double* yy, double* yd) { man_rhs<D>(P.man, yy, yd, rhoesrc, rYsrc); });
}

// 1-D form (ReactorRK64.cpp:22-164): T carried from the vector.
template <int D>
__global__ void __launch_bounds__(64) man_react_1d_kernel(ReactParamsDev P, double* __restrict__ yv,
                                                          const double* __restrict__ ysrc,
                                                          double* __restrict__ rX,
                                                          const double* __restrict__ rXsrc, double dt,
                                                          int ncells, CellStats st) {
  constexpr int NS = D + 1, n = D + 2;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
    double y[n], src[NS];
    for (int k = 0; k < n; ++k) y[k] = yv[(long long)c * n + k];
    for (int k = 0; k < NS; ++k) src[k] = ysrc[(long long)c * NS + k];
    const double rhoesrc = rXsrc[c];
    const int r = man_rk64_cell<D>(P, y, rhoesrc, src, dt);
    for (int k = 0; k < n; ++k) yv[(long long)c * n + k] = y[k];
    rX[c] = rX[c] + dt * rhoesrc;
    const int cnt[5] = {0, 0, 0, 0, 0};
    store_counters(st, c, r, cnt);
  }
}

// Box form (ReactorRK64.cpp:166-320).
template <int D>
__global__ void __launch_bounds__(64) man_react_box_kernel(ReactParamsDev P, BoxArgs b, double dt, int ncells,
                                                           CellStats st) {
  constexpr int NS = D + 1;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
    const int i = b.lo[0] + c % b.len[0];
    const int j = b.lo[1] + (c / b.len[0]) % b.len[1];
    const int k = b.lo[2] + c / (b.len[0] * b.len[1]);
    double y[NS + 1], src[NS], Y[D];
    for (int s = 0; s < NS; ++s) {
      y[s] = b.rY(i, j, k, s);
      src[s] = b.rYsrc(i, j, k, s);
    }
    double rinv = 1.0 / y[D];
    for (int d = 0; d < D; ++d) Y[d] = y[d] * rinv;
    y[NS] = man_temperature<D>(P.man, Y);
    const double rhoe = b.rE(i, j, k, 0), rhoesrc = b.rEsrc(i, j, k, 0);
    const int r = man_rk64_cell<D>(P, y, rhoesrc, src, dt);
    for (int s = 0; s < NS; ++s) b.rY(i, j, k, s) = y[s];
    rinv = 1.0 / y[D];
    for (int d = 0; d < D; ++d) Y[d] = y[d] * rinv;
    b.rE(i, j, k, 0) = rhoe + dt * rhoesrc;
    b.T(i, j, k, 0) = man_temperature<D>(P.man, Y);
    if (b.FC.p) b.FC(i, j, k, 0) = (double)r;
    const int cnt[5] = {0, 0, 0, 0, 0};
    store_counters(st, c, r, cnt);
  }
}

}  // namespace ppc
