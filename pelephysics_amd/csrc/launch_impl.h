// Kernel launch definitions, included by the per-(mechanism, integrator)
// instantiation units only.
#pragma once
#include <type_traits>
#include "launch.h"
#include "batched.h"
#include "wave_bdf.h"
#include "wave_adjoint.h"
#include "flatten.h"

namespace ppc {

// The batched forms run one lane per cell (batched.h); with PPC_WAVE_BATCHED=1
// table-driven mechanisms run them one wave per cell (wave_bdf.h) instead.
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

template <class M, int MODE>
hipError_t launch_1d_mode(const ReactParamsDev& P, double* y, const double* ysrc, double* rX,
                          const double* rXsrc, double dt, int ncells, CellStats st, int block, hipStream_t s) {
  const int grid = lane_grid<M, MODE>(P, ncells, block);
  if (grid < 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL((react_1d_kernel<M, MODE>), dim3(grid), dim3(block), 0, s, P, y, ysrc, rX, rXsrc, dt,
                     ncells, st);
  return hipGetLastError();
}

template <class M, int MODE>
hipError_t launch_box_mode(const ReactParamsDev& P, const BoxArgs& b, double dt, int ncells, CellStats st,
                           int block, hipStream_t s) {
  const int grid = lane_grid<M, MODE>(P, ncells, block);
  if (grid < 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL((react_box_kernel<M, MODE>), dim3(grid), dim3(block), 0, s, P, b, dt, ncells, st);
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

template <class M, int MODE>
hipError_t batched_1d_mode(const ReactParamsDev& P, double* y, const double* ysrc, double* rX, const double* rXsrc,
                           double dt, QueueArgs q, CellStats st, hipStream_t s) {
  using LU_T = typename std::conditional<MODE == MODE_BATCHED_MIXED, float, double>::type;
#if PPC_WAVE_BATCHED
  using Tab = typename wave_tables<M>::type;
  if constexpr (!std::is_void<Tab>::value) {
    auto k = wave_bdf_kernel<Tab, LU_T, WaveIO1D>;
    const int grid = resident_blocks(k, q.ncells * 64);
    if (grid < 0) return hipErrorInvalidDevice;
    WaveIO1D io{y, ysrc, rX, rXsrc};
    hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, s, P, io, dt, q, st);
  } else
#endif
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
#if PPC_WAVE_BATCHED
  using Tab = typename wave_tables<M>::type;
  if constexpr (!std::is_void<Tab>::value) {
    auto k = wave_bdf_kernel<Tab, LU_T, WaveIOBox>;
    const int grid = resident_blocks(k, q.ncells * 64);
    if (grid < 0) return hipErrorInvalidDevice;
    WaveIOBox io{b};
    hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, s, P, io, dt, q, st);
  } else
#endif
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
  if constexpr (!std::is_void<Tab>::value && M::NEQ <= 64 && M::NR <= 128) {
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

}  // namespace ppc

#define PPC_INSTANTIATE_FLATTEN(TYPE, MODE)                                                        \
  namespace ppc {                                                                                 \
  template hipError_t launch_flatten<TYPE>(const FlatArgs&, int, hipStream_t);                   \
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
