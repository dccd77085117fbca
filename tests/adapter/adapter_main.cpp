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
