// Test-only restatement of the interface include/ReactorHIP.H is written
// against: pele::physics::reactions::ReactorBase (Source/Reactions/
// ReactorBase.H:14-106, CPU build: no stream arguments), the Factory's
// self-registration (Source/Factory.H:60-162) and the few AMReX types and
// calls the adapter uses (amrex::Box, Array4, lbound/ubound, ParmParse,
// Abort).  It lets a CPU test compile the adapter, check every override
// against the reference signatures and exercise registration and error
// handling; it is not an AMReX or PelePhysics build.
#pragma once
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace amrex {
using Real = double;
struct Dim3 {
  int x, y, z;
};
struct Box {
  Dim3 lo, hi;
  long numPts() const { return (long)(hi.x - lo.x + 1) * (hi.y - lo.y + 1) * (hi.z - lo.z + 1); }
};
inline Dim3 lbound(const Box& b) { return b.lo; }
inline Dim3 ubound(const Box& b) { return b.hi; }
template <class T>
struct Array4 {
  T* p;
  long long jstride, kstride, nstride;
  Dim3 begin, end;
  int ncomp;
};
// ParmParse("ode").query(key, value): values from a process-wide table
struct ParmParse {
  explicit ParmParse(const std::string& prefix) : pre(prefix) {}
  static std::map<std::string, std::string>& table() {
    static std::map<std::string, std::string> t;
    return t;
  }
  template <class T>
  void query(const char* key, T& v) const {
    auto it = table().find(pre + "." + key);
    if (it == table().end()) return;
    if constexpr (std::is_same_v<T, std::string>) {
      v = it->second;
    } else if constexpr (std::is_same_v<T, bool>) {
      v = std::stoi(it->second) != 0;
    } else if constexpr (std::is_integral_v<T>) {
      v = (T)std::stoll(it->second);
    } else {
      v = (T)std::stod(it->second);
    }
  }
  std::string pre;
};
[[noreturn]] inline void Abort(const std::string& msg) { throw std::runtime_error(msg); }
template <class T, int N>
struct GpuArray {
  T arr[N];
  T& operator[](int i) { return arr[i]; }
  const T& operator[](int i) const { return arr[i]; }
};
namespace Gpu {
// AMReX_GpuDevice.H: device -> host copy (a plain copy on a CPU build)
inline void dtoh_memcpy(void* dst, const void* src, std::size_t n) { std::memcpy(dst, src, n); }
}  // namespace Gpu
}  // namespace amrex

// the species count of the compiled mechanism (mechanism.H); LiDryer here
#ifndef NUM_SPECIES
#define NUM_SPECIES 9
#endif

#define AMREX_ALWAYS_ASSERT(x) \
  do {                         \
    if (!(x)) amrex::Abort("assertion failed: " #x); \
  } while (0)

namespace pele::physics {

// The EOS parameter types the Manifold path of the adapter reads
// (Source/Utility/BlackBoxFunction/BlackBoxFunction.H:9-24, Table.H:11-17,
// Source/Eos/EosParams.H:15-50) and the build's EOS choice (PhysicsType).
enum class BlackBoxModel { TABLE = 0, NEURAL_NET = 1 };
struct BlackBoxFunctionData {
  BlackBoxModel bbmodel;
  static constexpr int len_str = 64;
  char model_name[len_str + 1];
  int Ndim;
  int Nvar;
  int Nmanpar;
  char* dimnames;
  char* varnames;
};
struct NeuralNetFunctionData : BlackBoxFunctionData {  // NeuralNetHomerolled.H:14-19
  double* nnrdata;
  int* nnidata;
  char nn_filename[len_str + 1];
};
struct TabulatedFunctionData : BlackBoxFunctionData {
  int varSpacing;
  int* dimLengths;
  int* dimDataSpacing;
  double* grids;
  double* values;
};
namespace eos {
enum class density_lookup_type { linear, log, inverse };
struct Fuego {};
struct Manifold {};
template <class EOS>
struct EosParm {};
template <>
struct EosParm<Manifold> {
  const BlackBoxFunctionData* manf_data;
  double Pnom_cgs;
  density_lookup_type dens_lookup;
  bool compute_temperature{false};
  int idx_density{0};
  int idx_T{0};
  int idx_Wdot[NUM_SPECIES - 1];
  bool has_mani_src{true};
};
}  // namespace eos
struct PhysicsType {
#ifdef USE_MANIFOLD_EOS
  using eos_type = eos::Manifold;
#else
  using eos_type = eos::Fuego;
#endif
};

template <class Base, class... Args>
struct Factory {
  static std::unique_ptr<Base> create(const std::string& key, Args... args) {
    if (table().find(key) == table().end()) amrex::Abort("In " + Base::base_identifier() + " cannot find instance: " + key);
    return table().at(key)(std::forward<Args>(args)...);
  }
  static std::vector<std::string> names() {
    std::vector<std::string> v;
    for (const auto& it : table()) v.push_back(it.first);
    return v;
  }
  template <class T, class DeriveFrom = Base>
  struct Register : public DeriveFrom {
    friend T;
    static bool registered;
    static bool add_sub_type() {
      Factory::table()[T::identifier()] = [](Args... args) -> std::unique_ptr<Base> {
        return std::unique_ptr<Base>(new T(std::forward<Args>(args)...));
      };
      return true;
    }
    ~Register() override = default;

   private:
    Register() { (void)registered; }
  };
  virtual ~Factory() = default;
  friend Base;

 private:
  using CreatorFunc = std::unique_ptr<Base> (*)(Args...);
  static std::map<std::string, CreatorFunc>& table() {
    static std::map<std::string, CreatorFunc> tbl;
    return tbl;
  }
  Factory() = default;
};

template <class Base, class... Args>
template <class T, class DeriveFrom>
bool Factory<Base, Args...>::Register<T, DeriveFrom>::registered =
    Factory<Base, Args...>::Register<T, DeriveFrom>::add_sub_type();

namespace reactions {

class ReactorBase : public Factory<ReactorBase> {
 public:
  static std::string base_identifier() { return "ReactorBase"; }
  virtual int init(int reactor_type, int ncells) = 0;
  virtual void close() = 0;
  virtual void print_final_stats(void* sundials_mem) = 0;
  virtual int react(const amrex::Box& box, amrex::Array4<amrex::Real> const& rY_in,
                    amrex::Array4<amrex::Real> const& rYsrc_in, amrex::Array4<amrex::Real> const& T_in,
                    amrex::Array4<amrex::Real> const& rEner_in, amrex::Array4<amrex::Real> const& rEner_src_in,
                    amrex::Array4<amrex::Real> const& FC_in, amrex::Array4<int> const& mask, amrex::Real& dt_react,
                    amrex::Real& time) = 0;
  virtual int react(amrex::Real* rY_in, amrex::Real* rYsrc_in, amrex::Real* rX_in, amrex::Real* rX_src_in,
                    amrex::Real& dt_react, amrex::Real& time, int ncells) = 0;
  virtual void flatten(const amrex::Box& box, const int ncells, amrex::Array4<const amrex::Real> const& rhoY,
                       amrex::Array4<const amrex::Real> const& frcExt,
                       amrex::Array4<const amrex::Real> const& temperature,
                       amrex::Array4<const amrex::Real> const& rhoE, amrex::Array4<const amrex::Real> const& frcEExt,
                       amrex::Real* y_vect, amrex::Real* src_vect, amrex::Real* vect_energy,
                       amrex::Real* src_vect_energy) = 0;
  virtual void unflatten(const amrex::Box& box, const int ncells, amrex::Array4<amrex::Real> const& rhoY,
                         amrex::Array4<amrex::Real> const& temperature, amrex::Array4<amrex::Real> const& rhoE,
                         amrex::Array4<amrex::Real> const& frcEExt, amrex::Array4<amrex::Real> const& FC_in,
                         amrex::Real* y_vect, amrex::Real* vect_energy, long int* FCunt, amrex::Real dt) = 0;
  ~ReactorBase() override = default;

 public:
  // ReactorBase.cpp:6-31 (not virtual)
  void set_typ_vals_ode(const std::vector<amrex::Real>& v)
  {
    if ((int)v.size() != NUM_SPECIES + 1) throw std::runtime_error("typ_vals size");
    for (int i = 0; i < NUM_SPECIES + 1; ++i) m_typ_vals[i] = v[i];
  }

  // ReactorBase.H:86-97
  virtual void set_eos_parm(const eos::EosParm<PhysicsType::eos_type>* eosparm)
  {
    if (std::is_same_v<PhysicsType::eos_type, eos::Manifold>) {
      amrex::Abort("Manifold EOS only supported with ReactorRK64 and ReactorNull for now");
    }
    m_eosparm = eosparm;
  }

 protected:
  int verbose{0};
  amrex::GpuArray<amrex::Real, NUM_SPECIES + 1> m_typ_vals = {{0.0}};
  const eos::EosParm<PhysicsType::eos_type>* m_eosparm{nullptr};
};

}  // namespace reactions
}  // namespace pele::physics
