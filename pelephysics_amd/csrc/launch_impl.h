// Kernel launch definitions, included by the per-(mechanism, integrator)
// instantiation units only.
#pragma once
#include <type_traits>
#include "launch.h"
#include "batched.h"
#include "wave_bdf.h"
#include "wave_adjoint.h"
#include "group_bdf.h"
#include "flatten.h"

namespace ppc {

// Table-driven mechanism tables of a mechanism type (void for the generated
// straight-line mechanisms): the group and wave-per-cell kernels read them.
template <class M>
struct wave_tables {
  using type = void;
};
template <class T>
struct wave_tables<TabMech<T>> {
  using type = T;
};

// Blocks of the one-lane-per-cell kernels: one lane per cell, or, with a
// workspace, exactly the lanes it holds (the lanes then stride over the cells).
template <class M, int MODE>
int lane_grid(const ReactParamsDev& P, int ncells, int block) {
  if (integrator_ws_rows<M>(MODE) > 0) {
    if (!P.ws || P.ws_lanes < block || P.ws_lanes % block != 0) return -1;
    return (int)(P.ws_lanes / block);
  }
  return (ncells + block - 1) / block;
}

// The plain adaptive BDF is built at the compiler's occupancy and, for
// mechanisms without a workspace, also held to 2 and 4 waves per SIMD
// (P.adapt.lane_occ); the parity integrators at the compiler's choice only.
template <class M, int MODE>
hipError_t launch_1d_mode(const ReactParamsDev& P, double* y, const double* ysrc, double* rX,
                          const double* rXsrc, double dt, int ncells, CellStats st, int block, hipStream_t s) {
  const int grid = lane_grid<M, MODE>(P, ncells, block);
  if (grid < 0) return hipErrorInvalidValue;
  if constexpr (MODE == MODE_ADAPTIVE && !uses_ws<M>()) {
    auto k = P.adapt.lane_occ == 4 ? react_1d_kernel_held<M, MODE, 4>
                                   : (P.adapt.lane_occ == 2 ? react_1d_kernel_held<M, MODE, 2> : react_1d_kernel<M, MODE>);
    hipLaunchKernelGGL(k, dim3(grid), dim3(block), 0, s, P, y, ysrc, rX, rXsrc, dt, ncells, st);
  } else {
    hipLaunchKernelGGL((react_1d_kernel<M, MODE>), dim3(grid), dim3(block), 0, s, P, y, ysrc, rX, rXsrc, dt,
                       ncells, st);
  }
  return hipGetLastError();
}

template <class M, int MODE>
hipError_t launch_box_mode(const ReactParamsDev& P, const BoxArgs& b, double dt, int ncells, CellStats st,
                           int block, hipStream_t s) {
  const int grid = lane_grid<M, MODE>(P, ncells, block);
  if (grid < 0) return hipErrorInvalidValue;
  if constexpr (MODE == MODE_ADAPTIVE && !uses_ws<M>()) {
    auto k = P.adapt.lane_occ == 4 ? react_box_kernel_held<M, MODE, 4>
                                   : (P.adapt.lane_occ == 2 ? react_box_kernel_held<M, MODE, 2> : react_box_kernel<M, MODE>);
    hipLaunchKernelGGL(k, dim3(grid), dim3(block), 0, s, P, b, dt, ncells, st);
  } else {
    hipLaunchKernelGGL((react_box_kernel<M, MODE>), dim3(grid), dim3(block), 0, s, P, b, dt, ncells, st);
  }
  return hipGetLastError();
}

// Persistent launch: as many 64-lane blocks as can be resident at once (never
// more than `lanes` / 64 needs); lanes or waves take cells from the queue.
template <class K>
int resident_blocks(K kernel, long long lanes) {
  int dev = 0, ncu = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, 0) != hipSuccess) return -1;
  if (per_cu < 1) per_cu = 1;
  const long long need = (lanes + 63) / 64;
  const long long blocks = (long long)ncu * per_cu;
  return (int)(blocks < need ? blocks : need);
}

// Persistent grid of a batched launch: every resident block, at most
// q.max_blocks when set and at most the lanes the workspace holds (-1: no
// device, -2: workspace missing).
template <class M, int MODE, class K>
int batched_grid(K kernel, const ReactParamsDev& P, const QueueArgs& q) {
  int grid = resident_blocks(kernel, q.ncells);
  if (grid < 0) return -1;
  if (q.max_blocks > 0 && q.max_blocks < grid) grid = q.max_blocks;
  if (integrator_ws_rows<M>(MODE) > 0) {
    if (!P.ws || P.ws_lanes < 64 || P.ws_lanes % 64 != 0) return -2;
    if (P.ws_lanes / 64 < grid) grid = (int)(P.ws_lanes / 64);
  }
  return grid;
}

// The group kernel (group_bdf.h) with G lanes per cell: a persistent grid of
// every resident block, at most q.max_blocks and at most the blocks the
// workspace holds (group_ws_rows rows of 64 lanes per block).
template <class T, typename LU_T, int G>
hipError_t launch_group(const ReactParamsDev& P, double* y, const double* ysrc, double* rX, const double* rXsrc,
                        double dt, QueueArgs q, CellStats st, hipStream_t s) {
  static_assert(group_ws_doubles<T::NEQ, G, LU_T>() <= group_ws_doubles_max<T::NEQ, G>(), "workspace layout");
  // registers of the LU rows: above ~100 VGPRs hold the build to one wave per SIMD
  constexpr int lu_vgprs = (int)(((T::NEQ + G - 1) / G) * T::NEQ * sizeof(LU_T) / 4);
  auto k = group_bdf_kernel<T, LU_T, G, (lu_vgprs > 100 ? 1 : 2)>;
  int grid = resident_blocks(k, (long long)q.ncells * G);
  if (grid < 0) return hipErrorInvalidDevice;
  if (q.max_blocks > 0 && q.max_blocks < grid) grid = q.max_blocks;
  if (!P.ws || P.ws_lanes < 64 || P.ws_lanes % 64 != 0) return hipErrorInvalidValue;
  if (P.ws_lanes / 64 < grid) grid = (int)(P.ws_lanes / 64);
  IO1D io{y, ysrc, rX, rXsrc};
  hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, s, P, io, dt, q, st);
  return hipGetLastError();
}

template <class M, typename LU_T>
hipError_t batched_group(const ReactParamsDev& P, double* y, const double* ysrc, double* rX, const double* rXsrc,
                         double dt, QueueArgs q, CellStats st, hipStream_t s) {
  using Tab = typename wave_tables<M>::type;
  if constexpr (!std::is_void<Tab>::value) {
    switch (P.adapt.group) {
      case 8:
        if constexpr (group_ok<M, 8>()) return launch_group<Tab, LU_T, 8>(P, y, ysrc, rX, rXsrc, dt, q, st, s);
        break;
      case 4:
        if constexpr (group_ok<M, 16>()) return launch_group<Tab, LU_T, 16>(P, y, ysrc, rX, rXsrc, dt, q, st, s);
        break;
      case 2:
        if constexpr (group_ok<M, 32>()) return launch_group<Tab, LU_T, 32>(P, y, ysrc, rX, rXsrc, dt, q, st, s);
        break;
      case 1:
        if constexpr (group_ok<M, 64>()) return launch_group<Tab, LU_T, 64>(P, y, ysrc, rX, rXsrc, dt, q, st, s);
        break;
      default:
        break;
    }
  }
  return hipErrorInvalidValue;
}

template <class M, int MODE>
hipError_t batched_1d_mode(const ReactParamsDev& P, double* y, const double* ysrc, double* rX, const double* rXsrc,
                           double dt, QueueArgs q, CellStats st, hipStream_t s) {
  using LU_T = typename std::conditional<MODE == MODE_BATCHED_MIXED, float, double>::type;
  if (P.adapt.group > 0) return batched_group<M, LU_T>(P, y, ysrc, rX, rXsrc, dt, q, st, s);
  {
    auto k = P.adapt.lane_occ == 1 ? batched_kernel<M, LU_T, IO1D, 1> : batched_kernel<M, LU_T, IO1D, 2>;
    const int grid = batched_grid<M, MODE>(k, P, q);
    if (grid < 0) return grid == -2 ? hipErrorInvalidValue : hipErrorInvalidDevice;
    IO1D io{y, ysrc, rX, rXsrc};
    hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, s, P, io, dt, q, st);
  }
  return hipGetLastError();
}

template <class M, int MODE>
hipError_t batched_box_mode(const ReactParamsDev& P, const BoxArgs& b, double dt, QueueArgs q, CellStats st,
                            hipStream_t s) {
  using LU_T = typename std::conditional<MODE == MODE_BATCHED_MIXED, float, double>::type;
  {
    // built at 2 waves per SIMD only: the 1-wave build of the table-driven
    // drm19 box kernel returned failed cells, so 1-wave runs take the flat route
    auto k = batched_kernel<M, LU_T, IOBox, 2>;
    const int grid = batched_grid<M, MODE>(k, P, q);
    if (grid < 0) return grid == -2 ? hipErrorInvalidValue : hipErrorInvalidDevice;
    IOBox io{b};
    hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, s, P, io, dt, q, st);
  }
  return hipGetLastError();
}

template <class M>
hipError_t launch_adjoint(const ReactParamsDev& P, const AdjointArgs& a, hipStream_t s) {
  using Tab = typename wave_tables<M>::type;
  // one equation per lane, at most two reactions per lane
  // (not QSS mechanisms: the adjoint differentiates plain mass-action rates)
  if constexpr (!std::is_void<Tab>::value && !has_qss<M>::value && M::NEQ <= 64 && M::NR <= 128) {
    auto k = wave_adjoint_kernel<Tab>;
    const int grid = resident_blocks(k, (long long)a.ncells * 64);
    if (grid < 0) return hipErrorInvalidDevice;
    hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, s, P, a);
    return hipGetLastError();
  } else {
    return hipErrorNotSupported;
  }
}

template <class M>
hipError_t launch_flatten(const FlatArgs& a, int unflatten, hipStream_t s) {
  const long long ncells = (long long)a.b.len[0] * a.b.len[1] * a.b.len[2];
  if (ncells <= 0) return hipSuccess;
  const int grid = (int)((ncells + 255) / 256 < 8192 ? (ncells + 255) / 256 : 8192);
  if (unflatten)
    hipLaunchKernelGGL(unflatten_kernel<M>, dim3(grid), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(flatten_kernel<M>, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// productionRate (Mechanisms/drm19/mechanism.H:4421-5667), aJacobian
// (:10307-15357) or the net rates of progress q_f - q_r (comp_qfqr, :3600) of
// one cell per lane: sc[c*NS + k] (mol/m^3), T[c] (K); out[c*NS + k] = wdot_k
// (mol/m^3/s), out[c*n*n + ...] = the column-major Jacobian of cell c, or
// out[c*NR + j] = q_j (mol/m^3/s), written in place in the output array.
template <class M>
__global__ void __launch_bounds__(64) mech_eval_kernel(int what, const double* __restrict__ sc,
                                                       const double* __restrict__ T, int ncells, double* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  double s[M::NS];
  for (int k = 0; k < M::NS; ++k) s[k] = sc[(long long)c * M::NS + k];
  const double Tv = T[c];
  if (what == 0) {
    double w[M::NS];
    M::production_rate(w, s, Tv);
    for (int k = 0; k < M::NS; ++k) out[(long long)c * M::NS + k] = w[k];
  } else if (what == 3) {
    double* q = out + (long long)c * M::NR;  // written in place: NR can be large
    M::progress_rates(q, s, Tv);
  } else {
    double* J = out + (long long)c * M::NEQ * M::NEQ;
    if (what == 1)
      M::template jacobian<0>(J, s, Tv);
    else
      M::template jacobian<1>(J, s, Tv);
  }
}

template <class M>
hipError_t launch_mech_eval(int what, const double* sc, const double* T, int ncells, double* out, hipStream_t s) {
  if (what < 0 || what > 3) return hipErrorInvalidValue;
  if (ncells <= 0) return hipSuccess;
  hipLaunchKernelGGL(mech_eval_kernel<M>, dim3((ncells + 63) / 64), dim3(64), 0, s, what, sc, T, ncells, out);
  return hipGetLastError();
}

}  // namespace ppc

#define PPC_INSTANTIATE_FLATTEN(TYPE, MODE)                                                        \
  namespace ppc {                                                                                 \
  template hipError_t launch_flatten<TYPE>(const FlatArgs&, int, hipStream_t);                   \
  template hipError_t launch_mech_eval<TYPE>(int, const double*, const double*, int, double*,     \
                                             hipStream_t);                                        \
  }

#define PPC_INSTANTIATE_ADJOINT(TYPE, MODE)                                                           \
  namespace ppc {                                                                                    \
  template hipError_t launch_adjoint<TYPE>(const ReactParamsDev&, const AdjointArgs&, hipStream_t);  \
  }

#define PPC_INSTANTIATE_MODE(TYPE, MODE)                                                              \
  namespace ppc {                                                                                    \
  template hipError_t launch_1d_mode<TYPE, MODE>(const ReactParamsDev&, double*, const double*, \
                                                       double*, const double*, double, int, CellStats, \
                                                       int, hipStream_t);                             \
  template hipError_t launch_box_mode<TYPE, MODE>(const ReactParamsDev&, const BoxArgs&, double, \
                                                        int, CellStats, int, hipStream_t);            \
  }

#define PPC_INSTANTIATE_BATCHED(TYPE, MODE)                                                             \
  namespace ppc {                                                                                      \
  template hipError_t batched_1d_mode<TYPE, MODE>(const ReactParamsDev&, double*, const double*,  \
                                                        double*, const double*, double, QueueArgs,      \
                                                        CellStats, hipStream_t);                       \
  template hipError_t batched_box_mode<TYPE, MODE>(const ReactParamsDev&, const BoxArgs&, double,  \
                                                         QueueArgs, CellStats, hipStream_t);           \
  }
