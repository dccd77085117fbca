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

struct ReactParamsDev {
  int reactor_type;
  BDFParamsDev bdf;
  RK64ParamsDev rk64;
  AdaptParamsDev adapt;
  // per-lane matrix workspace of the larger mechanisms (lane_ws.h): ws_lanes
  // lanes x the integrator's rows, or null
  double* ws;
  long long ws_lanes;
};

}  // namespace ppc
