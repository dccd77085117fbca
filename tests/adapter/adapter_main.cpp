// Drives include/ReactorHIP.H against the interface restated in
// tests/adapter/ReactorBase.H: the factory holds the adapter's names, a bad
// reactor_type aborts with the reference's message, an unknown mechanism
// aborts, and without a GPU a valid configuration reports the HIP error.
#include <cstdio>
#include <string>

#include "ReactorHIP.H"

using pele::physics::reactions::ReactorBase;

int main() {
  for (const auto& n : ReactorBase::names()) std::printf("name %s\n", n.c_str());
  auto r = ReactorBase::create("ReactorBDF_HIP");
  try {
    r->init(3, 1);
    std::printf("no abort\n");
  } catch (const std::exception& e) {
    std::printf("abort reactor_type: %s\n", e.what());
  }
#ifdef USE_MANIFOLD_EOS
  // a Manifold build: mechanism "Manifold" with NUM_SPECIES - 1 parameters;
  // ReactorBDF refuses it at init (the library) and at set_eos_parm (the
  // reference's message); ReactorRK64 reaches the device
  try {
    r->init(2, 1);
  } catch (const std::exception& e) {
    std::printf("abort manifold bdf init: %s\n", e.what());
  }
  pele::physics::eos::EosParm<pele::physics::eos::Manifold> ep{};
  try {
    r->set_eos_parm(&ep);
  } catch (const std::exception& e) {
    std::printf("abort manifold bdf eos: %s\n", e.what());
  }
  try {
    ReactorBase::create("ReactorRK64_HIP")->init(2, 1);
    std::printf("manifold rk64 init ok\n");
  } catch (const std::exception& e) {
    std::printf("abort manifold rk64 init: %s\n", e.what());
  }
#endif
  amrex::ParmParse::table()["ode.mechanism"] = "nope";
  try {
    r->init(2, 1);
  } catch (const std::exception& e) {
    std::printf("abort mechanism: %s\n", e.what());
  }
  amrex::ParmParse::table()["ode.mechanism"] = "drm19";
  try {
    ReactorBase::create("ReactorNotThere");
  } catch (const std::exception& e) {
    std::printf("abort create: %s\n", e.what());
  }
  return 0;
}
