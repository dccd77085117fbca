// Common definitions for the gfx950 chemistry kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define PPC_HD __host__ __device__
#ifndef PPC_DI
#define PPC_DI __device__ __forceinline__
#endif
#ifndef PPC_DEV
#define PPC_DEV __device__
#endif
// Large per-cell routines (RHS, Jacobian, LU, T inversion).  Inlined: out of
// line (-DPPC_HEAVY="__device__ __noinline__") the batched drm19 box kernel
// measured 517 ms per 64^3 call against 431 ms inlined.
#ifndef PPC_HEAVY
#define PPC_HEAVY __device__ __forceinline__
#endif

namespace ppc {

// Universal gas constant in erg/(mol K) (Source/PhysicsConstants.H:11).
constexpr double RU = 8.31446261815324e7;
// ln(10) and log10(e).
constexpr double LN10 = 2.302585092994045684;
constexpr double LOG10E = 0.4342944819032518277;

}  // namespace ppc
