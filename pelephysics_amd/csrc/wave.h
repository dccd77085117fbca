// Wavefront primitives for kernels that give one cell to one 64-lane wave:
// broadcasts (v_readlane into scalar registers), DPP row reductions and
// ballots.  Every reduction leaves the bitwise-identical result in all lanes
// (pairwise sums are formed in commuted order on the two partners, and the
// four row totals are combined in scalar-broadcast form), so branches on
// them are uniform.  Call them with all 64 lanes active.
#pragma once
#include "ppc_common.h"

namespace ppc {
namespace wave {

PPC_DI int lane_id() { return (int)__lane_id(); }

// (hi, lo) 32-bit halves of a double, and back, in unsigned arithmetic
PPC_DI unsigned int lo32(double v) { return (unsigned int)(unsigned long long)__double_as_longlong(v); }
PPC_DI unsigned int hi32(double v) { return (unsigned int)((unsigned long long)__double_as_longlong(v) >> 32); }
PPC_DI double from32(unsigned int hi, unsigned int lo) {
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | (unsigned long long)lo));
}

PPC_DI int rdlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
PPC_DI float rdlane(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
PPC_DI double rdlane(double v, int l) {
  const unsigned int lo = (unsigned int)__builtin_amdgcn_readlane((int)lo32(v), l);
  const unsigned int hi = (unsigned int)__builtin_amdgcn_readlane((int)hi32(v), l);
  return from32(hi, lo);
}
// value of the first lane (all lanes must hold the same value already)
PPC_DI double uni(double v) {
  const unsigned int lo = (unsigned int)__builtin_amdgcn_readfirstlane((int)lo32(v));
  const unsigned int hi = (unsigned int)__builtin_amdgcn_readfirstlane((int)hi32(v));
  return from32(hi, lo);
}
PPC_DI int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// DPP lane exchange inside rows of 16 lanes
template <int CTRL>
PPC_DI double dpp(double v) {
  const unsigned int lo = (unsigned int)__builtin_amdgcn_update_dpp(0, (int)lo32(v), CTRL, 0xF, 0xF, false);
  const unsigned int hi = (unsigned int)__builtin_amdgcn_update_dpp(0, (int)hi32(v), CTRL, 0xF, 0xF, false);
  return from32(hi, lo);
}
template <int CTRL>
PPC_DI float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
enum : int { DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140 };

template <typename T>
PPC_DI T sum(T v) {
  v += dpp<DPP_XOR1>(v);
  v += dpp<DPP_XOR2>(v);
  v += dpp<DPP_HALF_MIRROR>(v);
  v += dpp<DPP_MIRROR>(v);
  return (rdlane(v, 0) + rdlane(v, 16)) + (rdlane(v, 32) + rdlane(v, 48));
}
template <typename T>
PPC_DI T max(T v) {
  v = fmax(v, dpp<DPP_XOR1>(v));
  v = fmax(v, dpp<DPP_XOR2>(v));
  v = fmax(v, dpp<DPP_HALF_MIRROR>(v));
  v = fmax(v, dpp<DPP_MIRROR>(v));
  return fmax(fmax(rdlane(v, 0), rdlane(v, 16)), fmax(rdlane(v, 32), rdlane(v, 48)));
}

PPC_DI unsigned long long ballot(bool p) { return __ballot(p); }
// lowest set lane of a ballot, or `none` for an empty ballot
PPC_DI int first_lane(unsigned long long m, int none) { return m ? __ffsll((long long)m) - 1 : none; }

}  // namespace wave
}  // namespace ppc
