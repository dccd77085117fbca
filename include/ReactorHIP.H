// Header-only adapter: registers the gfx950 integrators behind
// pele::physics::reactions::ReactorBase (Source/Reactions/ReactorBase.H:14-106)
// so PeleC / PeleLMeX select them by name, exactly like the reference's
// ReactorBDF / ReactorRK64 (Source/Factory.H:72-79,95-124,159-162).
//
// Include this header in one translation unit of an AMReX + PelePhysics build
// (it needs the reference's ReactorBase.H, AMReX and a HIP build) and link
// libpele_chem_hip.so.  Names registered:
//   "ReactorBDF_HIP"          -> PP_CHEM_BDF           (ReactorBDF semantics)
//   "ReactorRK64_HIP"         -> PP_CHEM_RK64          (ReactorRK64 semantics)
//   "ReactorBDFAdaptive_HIP"  -> PP_CHEM_BDF_ADAPTIVE  (adaptive FP64 BDF)
//   "ReactorBDFBatched_HIP"   -> PP_CHEM_BDF_ADAPTIVE_MIXED (batched, FP32 LU)
//   "ReactorCvode_HIP"        -> PP_CHEM_CVODE         (ReactorCvode's options and
//                                defaults on the FP64 adaptive BDF; cvode.max_order,
//                                cvode.max_substeps, ode.max_nls_iters and the
//                                typical values of set_typ_vals_ode)
//   "ReactorArkode_HIP"       -> PP_CHEM_ARKODE        (ReactorArkode's explicit RK
//                                tables and step controllers: ode.rk_method,
//                                ode.rk_controller, the typical values)
// The ode.* ParmParse keys of ReactorBDF::init / ReactorRK64::init are read
// with the same names; ode.rtol / ode.atol / ode.max_steps for the adaptive
// forms.  The mechanism name comes from ode.mechanism (default: the string of
// the compiled PelePhysics mechanism passed as PP_CHEM_MECHANISM).
// In a Manifold build (USE_MANIFOLD_EOS: Eos_Model=Manifold, Chemistry_Model=
// Null) the mechanism is "Manifold" with NUM_SPECIES - 1 parameters, and
// ReactorRK64_HIP takes the tabulated model through set_eos_parm, as
// ReactorRK64 does (ReactorRK64.H:123-128); the other integrators abort there
// with the reference's message (ReactorBase.H:86-97).
// With PP_CHEM_REFERENCE_NAMES defined (a build that leaves the reference's
// ReactorBDF.cpp / ReactorRK64.cpp / ReactorCvode.cpp out of PelePhysics),
// "ReactorBDF", "ReactorRK64", "ReactorCvode" and "ReactorArkode" are
// registered too, so an unchanged inputs file selects them.
#pragma once

#include <string>
#include <vector>

#include "ReactorBase.H"
#include "pele_chem_hip.h"

namespace pele::physics::reactions {

template <int MODE, const char* NAME>
class ReactorHIP : public ReactorBase::Register<ReactorHIP<MODE, NAME>>
{
public:
  static std::string identifier() { return NAME; }

  int init(int reactor_type, int /*ncells*/) override
  {
    amrex::ParmParse pp("ode");
    std::string mech =
#if defined(USE_MANIFOLD_EOS)
      "Manifold";
#elif defined(PP_CHEM_MECHANISM)
      PP_CHEM_MECHANISM;
#else
      "drm19";
#endif
    pp.query("mechanism", mech);
    if (pp_chem_config_init(&m_cfg, mech.c_str(), MODE, reactor_type) != 0) {
      amrex::Abort("ReactorHIP: mechanism not compiled into libpele_chem_hip");
    }
    if (mech == "Manifold") {
      m_cfg.manifold_dim = NUM_SPECIES - 1;  // Mechanisms/Null/mechanism.H:11
    }
    pp.query("verbose", this->verbose);
    pp.query("bdf_nonlinear_iters", m_cfg.bdf_nonlinear_iters);
    pp.query("bdf_nonlinear_tol", m_cfg.bdf_nonlinear_tol);
    pp.query("bdf_nsubsteps", m_cfg.bdf_nsubsteps);
    pp.query("bdf_gmres_restarts", m_cfg.bdf_gmres_restarts);
    pp.query("bdf_gmres_kspiters", m_cfg.bdf_gmres_kspiters);
    pp.query("bdf_gmres_tol", m_cfg.bdf_gmres_tol);
    pp.query("bdf_gmres_precond", m_cfg.bdf_gmres_precond);
    pp.query("bdf_scheme", m_cfg.bdf_scheme);
    pp.query("atol", m_cfg.rk64_atol);
    pp.query("rk64_nsubsteps_guess", m_cfg.rk64_nsubsteps_guess);
    pp.query("rk64_nsubsteps_min", m_cfg.rk64_nsubsteps_min);
    pp.query("rk64_nsubsteps_max", m_cfg.rk64_nsubsteps_max);
    pp.query("rtol", m_cfg.rtol);
    pp.query("atol", m_cfg.atol);
    pp.query("max_steps", m_cfg.max_steps);
    pp.query("batched_waves_per_cu", m_cfg.batched_waves_per_cu);  // autotuner output
    pp.query("batched_lu_batch", m_cfg.batched_lu_batch);
    pp.query("batched_lu_wait", m_cfg.batched_lu_wait);
    pp.query("batched_waves_per_simd", m_cfg.batched_waves_per_simd);
    pp.query("batched_cells_per_wave", m_cfg.batched_cells_per_wave);
    pp.query("batched_box_direct", m_cfg.batched_box_direct);
    pp.query("workspace_bytes", m_cfg.workspace_bytes);
    if constexpr (MODE == PP_CHEM_CVODE) {
      // ReactorCvode.cpp:20-31
      pp.query("max_nls_iters", m_cfg.max_nls_iters);
      amrex::ParmParse ppcv("cvode");
      ppcv.query("max_order", m_cfg.cvode_max_order);
      ppcv.query("max_substeps", m_cfg.max_steps);
    }
    if constexpr (MODE == PP_CHEM_ARKODE) {
      // ReactorArkode.cpp:12-18
      pp.query("rk_method", m_cfg.erk_method);
      pp.query("rk_controller", m_cfg.erk_controller);
    }
    bool clean = false;
    pp.query("clean_init_massfrac", clean);  // ReactorBDF.cpp:127, ReactorRK64.cpp:18
    m_cfg.clean_init_massfrac = clean ? 1 : 0;
#ifdef MOD_REACTOR
    m_cfg.mod_reactor = 1;
#endif
    m_mech = mech;
    m_cfg.mechanism = m_mech.c_str();
    if (m_h) {
      pp_chem_destroy(m_h);
      m_h = nullptr;
    }
    const int rc = pp_chem_create(&m_cfg, &m_h);
    if (rc != 0) {
      if (reactor_type != 1 && reactor_type != 2) {
        amrex::Abort("Wrong reactor type. Choose between 1 (e) or 2 (h).");  // ReactorTypes.H:12-19
      }
      amrex::Abort("ReactorHIP: pp_chem_create failed with code " + std::to_string(rc) +
                   (rc == PP_CHEM_EHIP ? " (no usable GPU)" : " (invalid ode.* configuration)"));
    }
    return 0;
  }

  void close() override
  {
    if (m_h) {
      pp_chem_destroy(m_h);
      m_h = nullptr;
    }
  }

  void print_final_stats(void* /*mem*/) override {}

  int react(
    const amrex::Box& box,
    amrex::Array4<amrex::Real> const& rY_in,
    amrex::Array4<amrex::Real> const& rYsrc_in,
    amrex::Array4<amrex::Real> const& T_in,
    amrex::Array4<amrex::Real> const& rEner_in,
    amrex::Array4<amrex::Real> const& rEner_src_in,
    amrex::Array4<amrex::Real> const& FC_in,
    amrex::Array4<int> const& /*mask*/,
    amrex::Real& dt_react,
    amrex::Real& time
#ifdef AMREX_USE_GPU
    ,
    amrex::gpuStream_t stream
#endif
    ) override
  {
    const auto lo = amrex::lbound(box);
    const auto hi = amrex::ubound(box);
    const int l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
    const pp_array4 a[6] = {view(rY_in), view(rYsrc_in), view(T_in),
                            view(rEner_in), view(rEner_src_in), view(FC_in)};
    void* s = nullptr;
#ifdef AMREX_USE_GPU
    s = (void*)stream;
#endif
    sync_typ_vals();
    const int r = pp_chem_react_box(
      m_h, l, h, &a[0], &a[1], &a[2], &a[3], &a[4], &a[5], nullptr, &dt_react, &time, s);
    if (r < 0) {
      amrex::Abort(std::string("ReactorHIP: ") + pp_chem_last_error(m_h));
    }
    return r;
  }

  int react(
    amrex::Real* rY_in,
    amrex::Real* rYsrc_in,
    amrex::Real* rX_in,
    amrex::Real* rX_src_in,
    amrex::Real& dt_react,
    amrex::Real& time,
    int ncells
#ifdef AMREX_USE_GPU
    ,
    amrex::gpuStream_t stream
#endif
    ) override
  {
    void* s = nullptr;
#ifdef AMREX_USE_GPU
    s = (void*)stream;
#endif
    sync_typ_vals();
    // the reference's 1-D overload receives host pointers on GPU builds
    const int r = pp_chem_react_1d(
      m_h, rY_in, rYsrc_in, rX_in, rX_src_in, &dt_react, &time, ncells, 0, s);
    if (r < 0) {
      amrex::Abort(std::string("ReactorHIP: ") + pp_chem_last_error(m_h));
    }
    return r;
  }

  void flatten(
    const amrex::Box& box,
    const int ncells,
    amrex::Array4<const amrex::Real> const& rhoY,
    amrex::Array4<const amrex::Real> const& frcExt,
    amrex::Array4<const amrex::Real> const& temperature,
    amrex::Array4<const amrex::Real> const& rhoE,
    amrex::Array4<const amrex::Real> const& frcEExt,
    amrex::Real* y_vect,
    amrex::Real* src_vect,
    amrex::Real* vect_energy,
    amrex::Real* src_vect_energy) override
  {
    AMREX_ALWAYS_ASSERT(ncells == box.numPts());
    const auto lo = amrex::lbound(box);
    const auto hi = amrex::ubound(box);
    const int l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
    const pp_array4 a[5] = {view(rhoY), view(frcExt), view(temperature), view(rhoE), view(frcEExt)};
    const int r = pp_chem_flatten(
      m_h, l, h, &a[0], &a[1], &a[2], &a[3], &a[4], y_vect, src_vect, vect_energy,
      src_vect_energy, current_stream());
    if (r < 0) {
      amrex::Abort(std::string("ReactorHIP: ") + pp_chem_last_error(m_h));
    }
  }

  void unflatten(
    const amrex::Box& box,
    const int ncells,
    amrex::Array4<amrex::Real> const& rhoY,
    amrex::Array4<amrex::Real> const& temperature,
    amrex::Array4<amrex::Real> const& rhoE,
    amrex::Array4<amrex::Real> const& frcEExt,
    amrex::Array4<amrex::Real> const& FC_in,
    amrex::Real* y_vect,
    amrex::Real* vect_energy,
    long int* FCunt,
    amrex::Real dt) override
  {
    static_assert(sizeof(long int) == sizeof(long long), "FCunt is read as 64-bit counts");
    AMREX_ALWAYS_ASSERT(ncells == box.numPts());
    const auto lo = amrex::lbound(box);
    const auto hi = amrex::ubound(box);
    const int l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
    const pp_array4 a[5] = {view(rhoY), view(temperature), view(rhoE), view(frcEExt), view(FC_in)};
    const int r = pp_chem_unflatten(
      m_h, l, h, &a[0], &a[1], &a[2], &a[3], &a[4], y_vect, vect_energy,
      reinterpret_cast<const long long*>(FCunt), dt, current_stream());
    if (r < 0) {
      amrex::Abort(std::string("ReactorHIP: ") + pp_chem_last_error(m_h));
    }
  }

#ifdef USE_MANIFOLD_EOS
  // ReactorRK64::set_eos_parm (ReactorRK64.H:123-128) with the tabulated
  // manifold model: the parameter block is PeleParams' device copy and the
  // table's arrays are pinned host memory (Table.H:47-139), so the block and
  // the table header are copied to the host and the arrays handed over as
  // they are.
  void set_eos_parm(
    const pele::physics::eos::EosParm<pele::physics::PhysicsType::eos_type>* eosparm) override
  {
    if constexpr (MODE != PP_CHEM_RK64) {
      amrex::Abort("Manifold EOS only supported with ReactorRK64 and ReactorNull for now");
    } else {
      pele::physics::eos::EosParm<pele::physics::eos::Manifold> ep;
      amrex::Gpu::dtoh_memcpy(&ep, eosparm, sizeof(ep));
      pele::physics::BlackBoxFunctionData bb;
      amrex::Gpu::dtoh_memcpy(&bb, ep.manf_data, sizeof(bb));
      pp_chem_manifold m{};
      pele::physics::TabulatedFunctionData td;
      // NNModel::pack (NeuralNetModelDef.H:283-305) of a CMLM net, in the
      // pinned host buffers of NeuralNetFunctionData: ints {1, inputs, then
      // per layer Linear {nout, nin}, LeakyReLU {}, BatchNorm1d {n}}, reals
      // the layers' parameters one after the other
      int nn_kind[7], nn_nout[7];
      if (bb.bbmodel == pele::physics::BlackBoxModel::NEURAL_NET) {
        pele::physics::NeuralNetFunctionData nd;
        amrex::Gpu::dtoh_memcpy(&nd, ep.manf_data, sizeof(nd));
        const int* id = nd.nnidata;
        if (!id || !nd.nnrdata || id[0] != 1) {
          amrex::Abort("ReactorHIP: Only CMLM nets supported for now without owning data");
        }
        id += 2;
        const int order[7] = {0, 1, 2, 0, 1, 2, 0};  // Linear, LeakyReLU, BatchNorm1d, ...
        int width = nd.Ndim;
        for (int l = 0; l < 7; ++l) {
          nn_kind[l] = order[l];
          if (order[l] == 0) {
            width = id[0];
            id += 2;
          } else if (order[l] == 2) {
            id += 1;
          }
          nn_nout[l] = width;
        }
        m.model = 1;
        m.nn_layers = 7;
        m.nn_kind = nn_kind;
        m.nn_nout = nn_nout;
        m.nn_params = nd.nnrdata;
        m.ndim = nd.Ndim;
        m.nvar = nd.Nvar;
      } else {
        amrex::Gpu::dtoh_memcpy(&td, ep.manf_data, sizeof(td));
        m.ndim = td.Ndim;
        m.dim_lengths = td.dimLengths;
        m.grids = td.grids;
        m.nvar = td.Nvar;
        m.values = td.values;
      }
      m.idx_density = ep.idx_density;
      m.density_lookup = static_cast<int>(ep.dens_lookup);
      m.idx_T = ep.idx_T;
      m.compute_temperature = ep.compute_temperature ? 1 : 0;
      for (int k = 0; k < 5; ++k) {
        m.idx_src[k] = k < m.ndim ? ep.idx_Wdot[k] : -1;
      }
      m.nominal_pressure = ep.Pnom_cgs;
      if (pp_chem_set_manifold(m_h, &m) != 0) {
        amrex::Abort(std::string("ReactorHIP: ") + pp_chem_last_error(m_h));
      }
      this->m_eosparm = eosparm;
    }
  }
#endif

  ~ReactorHIP() override { close(); }

private:
  // flatten/unflatten carry no stream argument: AMReX's current stream
  static void* current_stream()
  {
#ifdef AMREX_USE_GPU
    return (void*)amrex::Gpu::gpuStream();
#else
    return nullptr;
#endif
  }

  // ReactorCvode's absolute tolerances follow the typical values the base
  // keeps (ReactorBase::set_typ_vals_ode is not virtual): handed over before
  // a call whenever they changed
  void sync_typ_vals()
  {
    if constexpr (MODE == PP_CHEM_CVODE || MODE == PP_CHEM_ARKODE) {
      constexpr int n = NUM_SPECIES + 1;
      std::vector<double> v(n);
      for (int k = 0; k < n; ++k) {
        v[k] = this->m_typ_vals[k];
      }
      if (v != m_typ_sent) {
        if (pp_chem_set_typ_vals(m_h, v.data(), n) != 0) {
          amrex::Abort(std::string("ReactorHIP: ") + pp_chem_last_error(m_h));
        }
        m_typ_sent = v;
      }
    }
  }

  template <class T>
  static pp_array4 view(amrex::Array4<T> const& a)
  {
    pp_array4 v;
    v.p = const_cast<double*>(reinterpret_cast<const double*>(a.p));
    v.jstride = a.jstride;
    v.kstride = a.kstride;
    v.nstride = a.nstride;
    v.begin[0] = a.begin.x;
    v.begin[1] = a.begin.y;
    v.begin[2] = a.begin.z;
    v.end[0] = a.end.x;
    v.end[1] = a.end.y;
    v.end[2] = a.end.z;
    v.ncomp = a.ncomp;
    return v;
  }

  pp_chem_config m_cfg{};
  std::string m_mech;
  pp_chem* m_h = nullptr;
  std::vector<double> m_typ_sent;
};

inline constexpr char kReactorBDF_HIP[] = "ReactorBDF_HIP";
inline constexpr char kReactorRK64_HIP[] = "ReactorRK64_HIP";
inline constexpr char kReactorBDFAdaptive_HIP[] = "ReactorBDFAdaptive_HIP";
inline constexpr char kReactorBDFBatched_HIP[] = "ReactorBDFBatched_HIP";
inline constexpr char kReactorCvode_HIP[] = "ReactorCvode_HIP";
inline constexpr char kReactorArkode_HIP[] = "ReactorArkode_HIP";

template class ReactorHIP<PP_CHEM_BDF, kReactorBDF_HIP>;
template class ReactorHIP<PP_CHEM_RK64, kReactorRK64_HIP>;
template class ReactorHIP<PP_CHEM_BDF_ADAPTIVE, kReactorBDFAdaptive_HIP>;
template class ReactorHIP<PP_CHEM_BDF_ADAPTIVE_MIXED, kReactorBDFBatched_HIP>;
template class ReactorHIP<PP_CHEM_CVODE, kReactorCvode_HIP>;
template class ReactorHIP<PP_CHEM_ARKODE, kReactorArkode_HIP>;

// Odr-use of each registration flag: the names are in the factory table from
// static initialization on (explicit instantiation alone does not define the
// implicit constructors that would otherwise use them).
inline const bool kReactorHIPRegistered =
  ReactorBase::Register<ReactorHIP<PP_CHEM_BDF, kReactorBDF_HIP>>::registered &&
  ReactorBase::Register<ReactorHIP<PP_CHEM_RK64, kReactorRK64_HIP>>::registered &&
  ReactorBase::Register<ReactorHIP<PP_CHEM_BDF_ADAPTIVE, kReactorBDFAdaptive_HIP>>::registered &&
  ReactorBase::Register<ReactorHIP<PP_CHEM_BDF_ADAPTIVE_MIXED, kReactorBDFBatched_HIP>>::registered &&
  ReactorBase::Register<ReactorHIP<PP_CHEM_CVODE, kReactorCvode_HIP>>::registered &&
  ReactorBase::Register<ReactorHIP<PP_CHEM_ARKODE, kReactorArkode_HIP>>::registered;

#ifdef PP_CHEM_REFERENCE_NAMES
inline constexpr char kReactorBDF_ref[] = "ReactorBDF";
inline constexpr char kReactorRK64_ref[] = "ReactorRK64";
inline constexpr char kReactorCvode_ref[] = "ReactorCvode";
inline constexpr char kReactorArkode_ref[] = "ReactorArkode";
template class ReactorHIP<PP_CHEM_BDF, kReactorBDF_ref>;
template class ReactorHIP<PP_CHEM_RK64, kReactorRK64_ref>;
template class ReactorHIP<PP_CHEM_CVODE, kReactorCvode_ref>;
template class ReactorHIP<PP_CHEM_ARKODE, kReactorArkode_ref>;
inline const bool kReactorHIPReferenceNames =
  ReactorBase::Register<ReactorHIP<PP_CHEM_BDF, kReactorBDF_ref>>::registered &&
  ReactorBase::Register<ReactorHIP<PP_CHEM_RK64, kReactorRK64_ref>>::registered &&
  ReactorBase::Register<ReactorHIP<PP_CHEM_CVODE, kReactorCvode_ref>>::registered &&
  ReactorBase::Register<ReactorHIP<PP_CHEM_ARKODE, kReactorArkode_ref>>::registered;
#endif

} // namespace pele::physics::reactions
