// Per-mechanism launcher table: the C-ABI dispatches by mechanism name at run
// time.  Kernel instantiations live in one translation unit per (mechanism,
// integrator) pair (inst/<mech>_<mode>.hip) so they compile in parallel.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include "reactors_fwd.h"
#include "lane_ws.h"

namespace ppc {

struct MechLauncher {
  const char* name;
  int ns, nr;
  int batched_ok;  // the batched modes are available
  hipError_t (*react_1d)(int mode, const ReactParamsDev& P, double* y, const double* ysrc, double* rX,
                         const double* rXsrc, double dt, int ncells, CellStats st, int block,
                         hipStream_t s);
  hipError_t (*react_box)(int mode, const ReactParamsDev& P, const BoxArgs& b, double dt, int ncells,
                          CellStats st, int block, hipStream_t s);
  // batched (persistent, work-queue) forms; grid = number of 64-lane blocks
  hipError_t (*batched_1d)(int mode, const ReactParamsDev& P, double* y, const double* ysrc, double* rX,
                           const double* rXsrc, double dt, QueueArgs q, CellStats st, hipStream_t s);
  hipError_t (*batched_box)(int mode, const ReactParamsDev& P, const BoxArgs& b, double dt, QueueArgs q,
                            CellStats st, hipStream_t s);
  // discrete adjoint of the fixed-step BDF (table-driven mechanisms only)
  hipError_t (*adjoint_1d)(const ReactParamsDev& P, const AdjointArgs& a, hipStream_t s);
  // workspace rows per lane a mode needs (lane_ws.h); 0: none
  long long (*ws_rows)(int mode);
  // workspace rows per lane of the batched group kernel with cells_per_wave
  // cells per wave (group_bdf.h); <= 0: no such kernel for this mechanism
  long long (*group_ws_rows)(int cells_per_wave);
  // box <-> flat state (flatten.h): unflatten = 0 flattens, 1 unflattens
  hipError_t (*flatten)(const FlatArgs& a, int unflatten, hipStream_t s);
  // the mechanism functions over a batch of cells: what = 0 productionRate,
  // 1 / 2 aJacobian with consP = 0 / 1, 3 net rates of progress (SI units, as
  // the reference's)
  hipError_t (*mech_eval)(int what, const double* sc, const double* T, int ncells, double* out, hipStream_t s);
};

template <class M>
hipError_t launch_mech_eval(int what, const double* sc, const double* T, int ncells, double* out, hipStream_t s);

template <class M>
hipError_t launch_flatten(const FlatArgs& a, int unflatten, hipStream_t s);

template <class M>
hipError_t launch_adjoint(const ReactParamsDev& P, const AdjointArgs& a, hipStream_t s);

template <class M, int MODE>
hipError_t launch_1d_mode(const ReactParamsDev& P, double* y, const double* ysrc, double* rX,
                          const double* rXsrc, double dt, int ncells, CellStats st, int block, hipStream_t s);
template <class M, int MODE>
hipError_t launch_box_mode(const ReactParamsDev& P, const BoxArgs& b, double dt, int ncells, CellStats st,
                           int block, hipStream_t s);

template <class M, int MODE>
hipError_t batched_1d_mode(const ReactParamsDev& P, double* y, const double* ysrc, double* rX, const double* rXsrc,
                           double dt, QueueArgs q, CellStats st, hipStream_t s);
template <class M, int MODE>
hipError_t batched_box_mode(const ReactParamsDev& P, const BoxArgs& b, double dt, QueueArgs q, CellStats st,
                            hipStream_t s);

template <class M>
hipError_t launch_batched_1d(int mode, const ReactParamsDev& P, double* y, const double* ysrc, double* rX,
                             const double* rXsrc, double dt, QueueArgs q, CellStats st, hipStream_t s) {
  if (mode == MODE_BATCHED_MIXED) return batched_1d_mode<M, MODE_BATCHED_MIXED>(P, y, ysrc, rX, rXsrc, dt, q, st, s);
  if (mode == MODE_BATCHED_FP64) return batched_1d_mode<M, MODE_BATCHED_FP64>(P, y, ysrc, rX, rXsrc, dt, q, st, s);
  return hipErrorInvalidValue;
}

template <class M>
hipError_t launch_batched_box(int mode, const ReactParamsDev& P, const BoxArgs& b, double dt, QueueArgs q,
                              CellStats st, hipStream_t s) {
  if (mode == MODE_BATCHED_MIXED) return batched_box_mode<M, MODE_BATCHED_MIXED>(P, b, dt, q, st, s);
  if (mode == MODE_BATCHED_FP64) return batched_box_mode<M, MODE_BATCHED_FP64>(P, b, dt, q, st, s);
  return hipErrorInvalidValue;
}

template <class M>
hipError_t launch_1d(int mode, const ReactParamsDev& P, double* y, const double* ysrc, double* rX,
                     const double* rXsrc, double dt, int ncells, CellStats st, int block, hipStream_t s) {
  switch (mode) {
    case MODE_BDF: return launch_1d_mode<M, MODE_BDF>(P, y, ysrc, rX, rXsrc, dt, ncells, st, block, s);
    case MODE_RK64: return launch_1d_mode<M, MODE_RK64>(P, y, ysrc, rX, rXsrc, dt, ncells, st, block, s);
    case MODE_ADAPTIVE: return launch_1d_mode<M, MODE_ADAPTIVE>(P, y, ysrc, rX, rXsrc, dt, ncells, st, block, s);
    case MODE_ERK: return launch_1d_mode<M, MODE_ERK>(P, y, ysrc, rX, rXsrc, dt, ncells, st, block, s);
    default: return hipErrorInvalidValue;
  }
}

template <class M>
hipError_t launch_box(int mode, const ReactParamsDev& P, const BoxArgs& b, double dt, int ncells, CellStats st,
                      int block, hipStream_t s) {
  switch (mode) {
    case MODE_BDF: return launch_box_mode<M, MODE_BDF>(P, b, dt, ncells, st, block, s);
    case MODE_RK64: return launch_box_mode<M, MODE_RK64>(P, b, dt, ncells, st, block, s);
    case MODE_ADAPTIVE: return launch_box_mode<M, MODE_ADAPTIVE>(P, b, dt, ncells, st, block, s);
    case MODE_ERK: return launch_box_mode<M, MODE_ERK>(P, b, dt, ncells, st, block, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ppc


#define PPC_DEFINE_LAUNCHER(NAME, TYPE)                                                        \
  namespace ppc {                                                                              \
  const MechLauncher* launcher_##NAME() {                                                      \
    static const MechLauncher L = {#NAME, TYPE::NS, TYPE::NR,                                  \
                                   1,                                                          \
                                   &launch_1d<TYPE>, &launch_box<TYPE>,                        \
                                   &launch_batched_1d<TYPE>, &launch_batched_box<TYPE>,        \
                                   &launch_adjoint<TYPE>, &integrator_ws_rows<TYPE>,           \
                                   &group_ws_rows<TYPE>,                                       \
                                   &launch_flatten<TYPE>, &launch_mech_eval<TYPE>};            \
    return &L;                                                                                 \
  }                                                                                            \
  }
