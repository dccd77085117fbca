// Common definitions for the gfx950 chemistry kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define PPC_HD __host__ __device__
#define PPC_DI __device__ __forceinline__
#ifndef PPC_DEV
#define PPC_DEV __device__
#endif

namespace ppc {

// Universal gas constant in erg/(mol K) (Source/PhysicsConstants.H:11).
constexpr double RU = 8.31446261815324e7;
// ln(10) and log10(e).
constexpr double LN10 = 2.302585092994045684;
constexpr double LOG10E = 0.4342944819032518277;

}  // namespace ppc
