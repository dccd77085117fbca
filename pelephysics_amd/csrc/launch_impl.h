// Kernel launch definitions, included by the per-(mechanism, integrator)
// instantiation units only.
#pragma once
#include "launch.h"
#include "reactors.h"

namespace ppc {

template <class M, int MODE>
hipError_t launch_1d_mode(const ReactParamsDev& P, double* y, const double* ysrc, double* rX,
                          const double* rXsrc, double dt, int ncells, CellStats st, int block, hipStream_t s) {
  const dim3 grid((ncells + block - 1) / block), blk(block);
  hipLaunchKernelGGL((react_1d_kernel<M, MODE>), grid, blk, 0, s, P, y, ysrc, rX, rXsrc, dt, ncells, st);
  return hipGetLastError();
}

template <class M, int MODE>
hipError_t launch_box_mode(const ReactParamsDev& P, const BoxArgs& b, double dt, int ncells, CellStats st,
                           int block, hipStream_t s) {
  const dim3 grid((ncells + block - 1) / block), blk(block);
  hipLaunchKernelGGL((react_box_kernel<M, MODE>), grid, blk, 0, s, P, b, dt, ncells, st);
  return hipGetLastError();
}

}  // namespace ppc

#define PPC_INSTANTIATE_MODE(MECH, MODE)                                                              \
  namespace ppc {                                                                                    \
  template hipError_t launch_1d_mode<mech::MECH, MODE>(const ReactParamsDev&, double*, const double*, \
                                                       double*, const double*, double, int, CellStats, \
                                                       int, hipStream_t);                             \
  template hipError_t launch_box_mode<mech::MECH, MODE>(const ReactParamsDev&, const BoxArgs&, double, \
                                                        int, CellStats, int, hipStream_t);            \
  }
