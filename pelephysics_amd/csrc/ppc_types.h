// Parameter blocks passed by value to the kernels.
#pragma once

namespace ppc {

struct BDFParamsDev {  // ReactorBDF ode.bdf_* (Source/Reactions/ReactorBDF.H:118-128)
  int nsubsteps, nonlinear_iters, gmres_restarts, gmres_kspiters, gmres_precond, scheme;
  double nonlin_tol, gmres_tol;
};

struct RK64ParamsDev {  // ReactorRK64 ode.* (Source/Reactions/ReactorRK64.H:131-134)
  double atol;
  int nsub_guess, nsub_min, nsub_max;
};

struct AdaptParamsDev {  // adaptive BDF tolerances
  double rtol, atol, first_step;
  int max_steps;
  int lu_batch, lu_wait;  // batched forms: wave-shared Jacobian/LU service (0: defaults 64, 32)
  int lane_occ;           // batched forms: kernel build held to 1 or 2 waves per SIMD (0: 2)
  int group;              // batched forms: cells per wave of the group kernel (8, 4, 2, 1); 0: one cell per lane
  // CVODE-style options of the plain FP64 form (0 / null: SciPy's behaviour)
  int max_order;          // order cap 1..5 (cvode.max_order)
  int newton_maxiter;     // Newton iterations per attempt (ode.max_nls_iters)
  int jac_age;            // re-evaluate the Jacobian after this many steps
  int fc_rhs;             // box FC and counters report RHS evaluations (ReactorCvode), not steps
  const double* atolv;    // per-component absolute tolerances atol x typical value (device, NEQ)
};

// A tabulated manifold model (TabulatedFunctionData, Source/Utility/
// BlackBoxFunction/Table.H:11-17, with the Manifold EOS parameters of
// Source/Eos/EosParams.H:39-49): ndim table dimensions (the manifold
// parameters), nvar tabulated variables on the tensor grid, first dimension
// fastest.
constexpr int MAN_MAXD = 5;  // MAXD_TABLE (Table.H:4)
constexpr int MAN_NN_MAXL = 16;   // network layers
constexpr int MAN_NN_MAXW = 100;  // network width bound (MAXD_NETWORK)
struct ManifoldDev {
  int ndim, nvar;
  int dim_len[MAN_MAXD];     // grid points per dimension (>= 2)
  int dim_off[MAN_MAXD];     // offset of each dimension's grid in grids
  int dim_stride[MAN_MAXD];  // values between neighbours along each dimension
  int var_stride;            // values per variable
  const double* grids;       // device, sum of dim_len
  const double* values;      // device, nvar x var_stride
  int idx_density;           // variable holding RHO, lnRHO or invRHO
  int dens_lookup;           // 0 linear, 1 log, 2 inverse (density_lookup_type)
  int idx_T;                 // variable holding T
  int compute_T;             // 1: T from the table; 0: T = 1
  int idx_src[MAN_MAXD];     // source term of each manifold parameter, -1: none (zero)
  double pnom;               // nominal pressure, dyn/cm^2
  // model 1: a network in place of the table (layers in order; kinds 0
  // Linear, 1 LeakyReLU, 2 BatchNorm1d; their parameters at nn + nn_off[l])
  int model;
  int nn_layers;
  int nn_kind[MAN_NN_MAXL], nn_nin[MAN_NN_MAXL], nn_nout[MAN_NN_MAXL], nn_off[MAN_NN_MAXL];
  const double* nn;
};

struct ErkParamsDev {  // ReactorArkode: the table of ode.rk_method and ode.rk_controller
  const double* tab;   // device: c[s], A[s*s] (row-major), b[s], d[s]
  int s, q, p, fsal;   // stages, order, embedding order, first same as last
  int controller;      // 0 PID, 1 PI, 2 I, 3 explicit Gustafsson
};

struct ReactParamsDev {
  int reactor_type;
  BDFParamsDev bdf;
  RK64ParamsDev rk64;
  AdaptParamsDev adapt;  // tolerances, step limit and typical values also of ReactorArkode
  ErkParamsDev erk;
  // per-lane matrix workspace of the larger mechanisms (lane_ws.h): ws_lanes
  // lanes x the integrator's rows, or null
  double* ws;
  long long ws_lanes;
  // the manifold model of the Manifold EOS (null values: none set)
  ManifoldDev man;
};

}  // namespace ppc
