/*
 * pele_chem_hip: C-ABI of the MI355X (gfx950) chemistry integrators.
 *
 * The entry points replace the GPU work of the reference reactors behind
 * pele::physics::reactions::ReactorBase (Source/Reactions/ReactorBase.H:14-106):
 *
 *   pp_chem_create / pp_chem_config_init
 *       <- ReactorBase::create(name) + ReactorX::init(reactor_type, ncells)
 *          (Source/Factory.H:72-79, ReactorBDF.cpp:112-130, ReactorRK64.cpp:6-20);
 *          the ode.* ParmParse keys become config fields.
 *   pp_chem_react_box
 *       <- ReactorX::react(const Box&, Array4 rY, rYsrc, T, rE, rEsrc, FC, mask,
 *          Real& dt_react, Real& time, gpuStream_t)
 *          (ReactorBase.H:25-40; ReactorBDF.cpp:314-494; ReactorRK64.cpp:166-320).
 *   pp_chem_react_1d
 *       <- ReactorX::react(Real* rY, Real* rYsrc, Real* rX, Real* rXsrc,
 *          Real& dt_react, Real& time, int ncells, gpuStream_t)
 *          (ReactorBase.H:42-54; ReactorBDF.cpp:133-311; ReactorRK64.cpp:22-164).
 *   pp_chem_destroy
 *       <- ReactorX::close() + delete (ReactorBase.H:21).
 *
 * Plain C types only: pointers, sizes, an Array4-shaped view struct.
 * Return values: >= 0 is int(average steps per cell) as the reference's
 * react() returns; < 0 is an error code (never abort); pp_chem_last_error()
 * gives the message.  A handle may be used from one host thread at a time.
 */
#ifndef PELE_CHEM_HIP_H
#define PELE_CHEM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* integrator selection (string names of ReactorBase::create) */
enum {
  PP_CHEM_BDF = 0,                 /* "ReactorBDF": fixed-step BDF, GMRES (parity) */
  PP_CHEM_RK64 = 1,                /* "ReactorRK64": explicit low-storage RK64 (parity) */
  PP_CHEM_BDF_ADAPTIVE = 2,        /* adaptive-order BDF with error control, one lane per cell, FP64 */
  PP_CHEM_BDF_ADAPTIVE_MIXED = 3,  /* batched adaptive BDF: persistent work queue, FP32 Newton-matrix
                                      LU, FP64 residuals, rates and error control */
  PP_CHEM_BDF_BATCHED_FP64 = 4,    /* batched adaptive BDF with an FP64 LU */
  PP_CHEM_CVODE = 5,               /* "ReactorCvode": the FP64 adaptive BDF (one lane per cell) with
                                      ReactorCvode's options and defaults (rtol 1e-6, atol 1e-10 x the
                                      typical values, order <= cvode.max_order = 2, ode.max_nls_iters
                                      = 3 Newton iterations, Jacobian age <= 100 steps,
                                      cvode.max_substeps = 10000); FC and the return value count RHS
                                      evaluations (ReactorCvode.cpp:1365-1372).  Not SUNDIALS' CVODE:
                                      the same option set on this library's NDF/BDF integrator. */
  PP_CHEM_ARKODE = 6               /* "ReactorArkode": explicit Runge-Kutta with the table of
                                      ode.rk_method (erk_method; ReactorArkode.cpp:26-83, all but codes
                                      43 and 54) and the error-based step control of ode.rk_controller
                                      (erk_controller: 0 PID, 1 PI, 2 I, 3 explicit Gustafsson), rtol /
                                      atol x typical values, one cell per lane with its own step size;
                                      FC and the return value count RHS evaluations
                                      (ReactorArkode.cpp:239-260).  Not SUNDIALS' ARKODE: its tables and
                                      published adaptivity on this library's stepping. */
};

enum {
  PP_CHEM_OK = 0,
  PP_CHEM_EINVAL = -1,    /* bad argument or config (e.g. reactor_type not 1 or 2) */
  PP_CHEM_EMECH = -2,     /* unknown mechanism name */
  PP_CHEM_EHIP = -3,      /* HIP runtime error */
  PP_CHEM_EINTEGRATE = -4 /* a cell failed (NaN in GMRES, step limit, step too small) */
};

typedef struct {
  const char* mechanism; /* "LiDryer", "drm19", ...; "Manifold": the Manifold EOS (below) */
  int reactor_type;      /* 1: e (constant volume, rho e), 2: h (rho h) */
  int mode;              /* PP_CHEM_* integrator */
  int mod_reactor;       /* 1: time += dt_react on return (-DMOD_REACTOR) */
  /* ReactorBDF ode.bdf_* keys (ReactorBDF.H:118-128) */
  int bdf_nsubsteps, bdf_nonlinear_iters, bdf_gmres_restarts, bdf_gmres_kspiters;
  int bdf_gmres_precond, bdf_scheme;
  double bdf_nonlinear_tol, bdf_gmres_tol;
  /* ReactorRK64 ode.* keys (ReactorRK64.H:131-134) */
  double rk64_atol;
  int rk64_nsubsteps_guess, rk64_nsubsteps_min, rk64_nsubsteps_max;
  /* adaptive BDF */
  double rtol, atol, first_step;
  int max_steps;
  int block_size; /* threads per block, 0: default (64) */
  /* Upper bound on the device workspace that holds the per-cell matrices of
   * mechanisms with more than 31 species (Jacobian, LU, Krylov basis; too large
   * for a lane's private memory); 0: default (32 GiB).  Fewer resident cells
   * are integrated at once when it is reached. */
  long long workspace_bytes;
  /* ode.clean_init_massfrac (ReactorBDF.cpp:127, ReactorRK64.cpp:18): the
   * box <-> flat conversion clips rhoY to [0, rho] and re-applies the
   * increment conservatively (pp_chem_flatten / pp_chem_unflatten). */
  int clean_init_massfrac;
  /* Batched forms: persistent waves per compute unit (0: every wave that fits,
   * 8 at the kernels' 2 waves per SIMD); a tuning parameter
   * (python -m pelephysics_amd.autotune). */
  int batched_waves_per_cu;
  /* Batched forms: a wave factors the Newton matrices of its waiting lanes
   * together once batched_lu_batch lanes (1..64) wait or a request has waited
   * batched_lu_wait trips (0: 64 and 32); results do not depend on them. */
  int batched_lu_batch, batched_lu_wait;
  /* Batched forms, box call: 0 (default) runs box -> flat, the 1-D kernel,
   * flat -> box (same per-cell operations, faster kernel build); 1 runs the
   * box kernel directly. */
  int batched_box_direct;
  /* The kernel build a launch uses.  Batched forms: held to 2 waves per SIMD
   * (0 or 2, the default) or to 1 (more registers per lane, fewer spills;
   * 1-D kernel only, so not with batched_box_direct = 1).  The plain adaptive
   * BDF (PP_CHEM_BDF_ADAPTIVE): 0 (the compiler's choice), 2 or 4.  The
   * parity integrators: 0. */
  int batched_waves_per_simd;
  /* Batched forms, cells per 64-lane wavefront: 0 (default) integrates one
   * cell per lane (64 per wave); 8, 4, 2 or 1 give each cell a group of 8, 16,
   * 32 or 64 lanes that share its equations, reactions and LU rows, so its
   * state stays in registers (table-driven mechanism kernels only; the values
   * a mechanism supports depend on its size; box calls go through flat
   * state). */
  int batched_cells_per_wave;
  /* PP_CHEM_BDF_ADAPTIVE and PP_CHEM_CVODE (0 elsewhere): order cap 1..5
   * (cvode.max_order; 0: 5), Newton iterations per step attempt
   * (ode.max_nls_iters; 0: 4), and the Jacobian's largest age in steps before
   * it is re-evaluated at the predicted state (CVodeSetJacEvalFrequency; 0:
   * only when Newton fails to converge). */
  int cvode_max_order, max_nls_iters, jac_max_age;
  /* mechanism "Manifold" (the reference's Eos_Model=Manifold with
   * Chemistry_Model=Null, Mechanisms/Null/mechanism.H:11): the number of
   * manifold parameters D = MANIFOLD_DIM (1..5; pp_chem_config_init sets 1),
   * NUM_SPECIES = D + 1 (the parameters times rho, then rho).  ReactorRK64
   * only, as the reference propagates the Manifold EOS through ReactorRK64
   * alone (ReactorBase.H:84-97, ReactorRK64.H:123-128). */
  int manifold_dim;
  /* PP_CHEM_ARKODE (0 elsewhere): ode.rk_method (pp_chem_config_init: 40,
   * ARKODE_ZONNEVELD_5_3_4; unknown codes select it, as the reference does;
   * 43 and 54 are refused) and ode.rk_controller (0 PID). */
  int erk_method, erk_controller;
} pp_chem_config;

/* amrex::Array4<Real> layout: element (i,j,k,n) is at
 * p[(i-begin[0]) + (j-begin[1])*jstride + (k-begin[2])*kstride + n*nstride]. */
typedef struct {
  double* p;
  long long jstride, kstride, nstride;
  int begin[3];
  int end[3];
  int ncomp;
} pp_array4;

typedef struct {
  long long ncells, nsteps, nrhs, njac, nlu, nnewton, nreject;
  int max_steps_cell, min_steps_cell, nfailed;
  double kernel_ms; /* device time of the last integration kernel */
} pp_chem_stats;

typedef struct pp_chem pp_chem;

/* ReactorBase::set_typ_vals_ode (ReactorBase.cpp:6-31, read by
 * ReactorUtils.H:119-131): n = NUM_SPECIES + 1 typical values (species, then
 * T); ReactorCvode then uses the absolute tolerances atol x typ_vals[k]
 * (typ_vals[0] <= 0: the scalar atol again).  The other integrators ignore
 * them, as the reference's ReactorBDF and ReactorRK64 do. */
int pp_chem_set_typ_vals(pp_chem* h, const double* typ_vals, int n);

/* The Manifold EOS's parameters with a tabulated model, the data of
 * ReactorRK64::set_eos_parm (ReactorRK64.H:123-128): EosParm<Manifold>
 * (Source/Eos/EosParams.H:39-49) over TabulatedFunctionData
 * (Source/Utility/BlackBoxFunction/Table.H:11-17, the .ctb file layout of
 * read_table, Table.H:63-139).  Host arrays, copied by pp_chem_set_manifold. */
typedef struct {
  int ndim;                /* table dimensions = manifold parameters (must equal manifold_dim) */
  const int* dim_lengths;  /* ndim grid lengths, each >= 2 */
  const double* grids;     /* the ndim grids, concatenated, each strictly ascending */
  int nvar;                /* tabulated variables */
  const double* values;    /* nvar x prod(dim_lengths), first dimension fastest */
  int idx_density;         /* variable holding RHO (density_lookup 0), lnRHO (1) or invRHO (2) */
  int density_lookup;
  int idx_T;               /* variable holding T */
  int compute_temperature; /* manifold.compute_temperature: 1: T from the table, 0: T = 1 */
  int idx_src[5];          /* source of each manifold parameter (SRC_<name>), -1: none (zero) */
  double nominal_pressure; /* manifold.nominal_pressure_cgs */
  /* model 1: a network (manifold.model = NeuralNet, NeuralNetHomerolled.H,
   * NeuralNetModelDef.H, NeuralNetLayerDef.H) in place of the table: ndim
   * inputs, nvar outputs, dim_lengths / grids / values unused.  Layer l of
   * nn_layers (at most 16) is nn_kind[l] = 0 Linear (nn_nout[l] outputs;
   * parameters W[nout][nin] then b[nout]), 1 LeakyReLU (the negative slope)
   * or 2 BatchNorm1d (scale[n] then shift[n]), each layer narrower than 100;
   * nn_params holds the layers' parameters one after the other.  model 0
   * (the zero default) is the table. */
  int model;
  int nn_layers;
  const int* nn_kind;
  const int* nn_nout;
  const double* nn_params;
} pp_chem_manifold;

/* ReactorRK64::set_eos_parm for the Manifold EOS: the table every lookup of
 * the next calls interpolates in (multilinear; indices clamped to the grid,
 * weights to [0, 1], Table.H:183-219,470-545).  A handle of mechanism
 * "Manifold" returns PP_CHEM_EINVAL from its react calls until this is set;
 * other mechanisms refuse it.  flatten / unflatten, the adjoint and the
 * mechanism functions are not available with the Manifold EOS. */
int pp_chem_set_manifold(pp_chem* h, const pp_chem_manifold* m);

/* Fill cfg with the reference defaults for the given mechanism and mode. */
int pp_chem_config_init(pp_chem_config* cfg, const char* mechanism, int mode, int reactor_type);
int pp_chem_create(const pp_chem_config* cfg, pp_chem** out);
void pp_chem_destroy(pp_chem* h);
int pp_chem_num_species(const pp_chem* h);
/* Comma-separated list of compiled mechanisms; returns the needed length. */
int pp_chem_list_mechanisms(char* buf, int buflen);

/* Box form: device-resident Array4 views over [lo, hi] (inclusive), updated
 * in place (rY, T, rE, FC); rYsrc, rEsrc read; mask ignored as in the
 * reference (ReactorBDF.cpp:322).  FC may have p == NULL.
 * The batched modes take the cells from a work queue; when the previous call
 * on this handle had the same cell count and the same rY data pointer (the
 * same cells, e.g. the next chemistry step of the same box), the queue visits
 * them in the order of that call's per-cell cost, longest first.  Results do
 * not depend on the order. */
int pp_chem_react_box(pp_chem* h, const int lo[3], const int hi[3], const pp_array4* rY,
                      const pp_array4* rYsrc, const pp_array4* T, const pp_array4* rE,
                      const pp_array4* rEsrc, const pp_array4* FC, const pp_array4* mask,
                      double* dt_react, double* time, void* stream);

/* 1-D form: y[icell*(N+1)+k] (k = N is T), ysrc[icell*N+k], rX/rXsrc[icell];
 * on_device = 0: host pointers (copied in and out, as the reference does on
 * GPU builds, ReactorBDF.cpp:155-171,295-302); 1: device pointers.  The cost
 * order of the batched modes applies when the previous call had the same
 * ncells and the same y pointer (host or device): reuse the buffers. */
int pp_chem_react_1d(pp_chem* h, double* y, double* ysrc, double* rX, double* rXsrc,
                     double* dt_react, double* time, int ncells, int on_device, void* stream);

/* Box -> flat state, ReactorBase::flatten (ReactorBase.H:56-67; FlattenOps<YCOrder>,
 * ReactorUtils.H:160-220,291-325): y_vect (ncells x (N+1), T from the energy
 * with the box T as guess), src_vect (ncells x N), vect_energy and
 * src_vect_energy (ncells); all device pointers, cell index
 * (k-lo2)*len0*len1 + (j-lo1)*len0 + (i-lo0).  With clean_init_massfrac the
 * flat rhoY are clipped to [0, sum rhoY]. */
int pp_chem_flatten(pp_chem* h, const int lo[3], const int hi[3], const pp_array4* rY, const pp_array4* rYsrc,
                    const pp_array4* T, const pp_array4* rE, const pp_array4* rEsrc, double* y_vect,
                    double* src_vect, double* vect_energy, double* src_vect_energy, void* stream);
/* Flat -> box state, ReactorBase::unflatten (ReactorBase.H:69-80;
 * ReactorUtils.H:222-289,327-355): rY from y_vect (with clean_init_massfrac:
 * old rhoY plus the increment over its clipped value, scaled by
 * rho_old / rho_clipped), rE = vect_energy + dt * rEsrc, T from the energy with
 * the flat T as guess, FC = FCunt (the reference's long int counts). */
int pp_chem_unflatten(pp_chem* h, const int lo[3], const int hi[3], const pp_array4* rY, const pp_array4* T,
                      const pp_array4* rE, const pp_array4* rEsrc, const pp_array4* FC, const double* y_vect,
                      const double* vect_energy, const long long* FCunt, double dt, void* stream);

/* ncalls successive 1-D calls of dt_react on device arrays (a flow solver's
 * chemistry steps without host work in between): the same results as ncalls
 * calls of pp_chem_react_1d (rX advanced by dt * rXsrc each call).  One call's
 * device work (queue reset, cost ordering, integration kernel) is captured
 * once as a HIP graph and replayed; the graph is rebuilt when an array,
 * ncells or dt changes.  Statistics and per-cell counters are those of the
 * last call; kernel_ms covers all calls.  Returns PP_CHEM_EINTEGRATE when a
 * cell failed in any of the calls (its partly integrated state is what the
 * next call starts from).  No counterpart in the reference,
 * whose callers loop over ReactorBDF::react (ReactEval/main.cpp:475-686). */
int pp_chem_react_1d_chained(pp_chem* h, double* y, double* ysrc, double* rX, double* rXsrc,
                             double* dt_react, double* time, int ncells, int ncalls, void* stream);

/* The mechanism functions the integrators call, over a batch of cells on the
 * device (one cell per lane), in the reference's SI units:
 * pp_chem_production_rate replaces productionRate(wdot, sc, T)
 * (Mechanisms/drm19/mechanism.H:4421-5667): sc[c*N + k] in mol/m^3, T[c] in K,
 * wdot[c*N + k] in mol/m^3/s;
 * pp_chem_ajacobian replaces aJacobian(J, sc, T, consP)
 * (Mechanisms/drm19/mechanism.H:10307-15357): J[c*(N+1)^2 + k*(N+1) + m] =
 * d wdot_m / d c_k, column N = d / dT, row N = the temperature equation.
 * pp_chem_progress_rate gives the net rate of progress q_f - q_r of every
 * reaction (the rates comp_qfqr forms, Mechanisms/drm19/mechanism.H:3600,
 * and progressRateFR returns split, mechanism.cpp:120): q[c*NR + j].
 * Synchronous on `stream`; the handle selects the mechanism. */
int pp_chem_production_rate(pp_chem* h, const double* sc, const double* T, int ncells, double* wdot,
                            void* stream);
int pp_chem_ajacobian(pp_chem* h, const double* sc, const double* T, int ncells, int consP, double* J,
                      void* stream);
int pp_chem_progress_rate(pp_chem* h, const double* sc, const double* T, int ncells, double* q, void* stream);

/* pp_chem_react_1d on device arrays without any device allocation: no
 * counterpart in the reference.  For callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same integration and results as pp_chem_react_1d with on_device = 1
 * (the queue visits the cells in index order; the per-cell step counts are
 * reduced, the other per-cell counters are not kept).  Its queue counter,
 * step counts and integrator workspace live in one file-scope device array of
 * 64 MiB per GPU, held until the process ends: at most 16,777,088 cells per
 * call (larger ncells returns PP_CHEM_EINVAL); a mechanism / integrator whose
 * workspace needs more than the array holds runs fewer resident blocks.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (other streams, other handles, other host threads) race. */
int pp_chem_react_1d_noalloc(pp_chem* h, double* y, double* ysrc, double* rX, double* rXsrc, double* dt_react,
                             double* time, int ncells, void* stream);

int pp_chem_get_stats(const pp_chem* h, pp_chem_stats* out);
int pp_chem_num_reactions(const pp_chem* h);

/* Discrete adjoint of the fixed-step BDF integration: no counterpart in the
 * reference (it exposes no sensitivities); the map is that of
 * ReactorBDF::react (ReactorBDF.cpp:187-285) with the handle's ode.bdf_*
 * parameters (bdf_nonlinear_iters must be 1), called nmacro times with dt each
 * (1-D form, no sources), with every linear system solved exactly.
 * y: ncells x (N+1) YCOrder state (device), overwritten with the final state;
 * rX: ncells rho e or rho h (device); out: ncells x 3 (device) = T_final,
 * Y[spec_a], Y[spec_b]; grad: ncells x NR x 3 (device) = d out / d ln A_j,
 * A_j the pre-exponential of reaction j in mechanism order (for falloff
 * reactions the high- and low-pressure limits together).  The device time of
 * the kernel is in pp_chem_get_stats().kernel_ms. */
int pp_chem_adjoint_1d(pp_chem* h, double* y, const double* rX, int ncells, int nmacro, double dt,
                       int spec_a, int spec_b, double* out, double* grad, void* stream);
/* Doubles per resident cell of the per-cell matrix workspace the handle's
 * integrator uses (0: the matrices stay in private memory).  No counterpart in
 * the reference; for sizing pp_chem_config.workspace_bytes. */
long long pp_chem_workspace_rows(const pp_chem* h);
/* Per-cell accepted step counts of the last call (host copy). */
int pp_chem_copy_cell_steps(pp_chem* h, int* host_out, int ncells);
const char* pp_chem_last_error(const pp_chem* h);

#ifdef __cplusplus
}
#endif
#endif /* PELE_CHEM_HIP_H */
