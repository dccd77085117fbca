// Launcher table of the Manifold EOS (ReactorRK64 only, as the reference's
// ReactorBase::set_eos_parm allows; ReactorBase.H:84-97, ReactorRK64.H:123-128)
// for tables of 1 to 5 dimensions.
#include "manifold.h"
#include "launch.h"

namespace ppc {
namespace {

template <int D>
hipError_t man_1d(int mode, const ReactParamsDev& P, double* y, const double* ysrc, double* rX, const double* rXsrc,
                  double dt, int ncells, CellStats st, int block, hipStream_t s) {
  if (mode != MODE_RK64 || P.man.ndim != D || (!P.man.values && !P.man.nn) || block < 1) return hipErrorInvalidValue;
  const int grid = (ncells + block - 1) / block;
  hipLaunchKernelGGL(man_react_1d_kernel<D>, dim3(grid), dim3(block), 0, s, P, y, ysrc, rX, rXsrc, dt, ncells, st);
  return hipGetLastError();
}

template <int D>
hipError_t man_box(int mode, const ReactParamsDev& P, const BoxArgs& b, double dt, int ncells, CellStats st,
                   int block, hipStream_t s) {
  if (mode != MODE_RK64 || P.man.ndim != D || (!P.man.values && !P.man.nn) || block < 1) return hipErrorInvalidValue;
  const int grid = (ncells + block - 1) / block;
  hipLaunchKernelGGL(man_react_box_kernel<D>, dim3(grid), dim3(block), 0, s, P, b, dt, ncells, st);
  return hipGetLastError();
}

hipError_t no_batched_1d(int, const ReactParamsDev&, double*, const double*, double*, const double*, double,
                         QueueArgs, CellStats, hipStream_t) {
  return hipErrorInvalidValue;
}
hipError_t no_batched_box(int, const ReactParamsDev&, const BoxArgs&, double, QueueArgs, CellStats, hipStream_t) {
  return hipErrorInvalidValue;
}
hipError_t no_adjoint(const ReactParamsDev&, const AdjointArgs&, hipStream_t) { return hipErrorNotSupported; }
long long no_ws(int) { return 0; }
long long no_group(int) { return 0; }
hipError_t no_flatten(const FlatArgs&, int, hipStream_t) { return hipErrorNotSupported; }
hipError_t no_eval(int, const double*, const double*, int, double*, hipStream_t) { return hipErrorNotSupported; }

template <int D>
const MechLauncher* man_launcher() {
  static const MechLauncher L = {"Manifold", D + 1, 0, 0, &man_1d<D>, &man_box<D>, &no_batched_1d,
                                 &no_batched_box, &no_adjoint, &no_ws, &no_group, &no_flatten, &no_eval};
  return &L;
}

}  // namespace

const MechLauncher* manifold_launcher(int dim) {
  switch (dim) {
    case 1: return man_launcher<1>();
    case 2: return man_launcher<2>();
    case 3: return man_launcher<3>();
    case 4: return man_launcher<4>();
    case 5: return man_launcher<5>();
    default: return nullptr;
  }
}

}  // namespace ppc
