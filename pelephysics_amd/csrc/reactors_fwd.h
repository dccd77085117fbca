// Types shared by the kernels and the host launchers.
#pragma once
#include "ppc_common.h"
#include "ppc_types.h"

namespace ppc {

enum : int { MODE_BDF = 0, MODE_RK64 = 1, MODE_ADAPTIVE = 2, MODE_ADAPTIVE_MIXED = 3 };

// ------------------------------------------------------------------ I/O --
struct A4 {  // amrex::Array4 layout
  double* p;
  long long js, ks, ns;
  int bx, by, bz;
  PPC_DI double& operator()(int i, int j, int k, int n) const {
    return p[(long long)(i - bx) + (long long)(j - by) * js + (long long)(k - bz) * ks + (long long)n * ns];
  }
};

struct BoxArgs {
  int lo[3], len[3];
  A4 rY, rYsrc, T, rE, rEsrc, FC;
};

struct CellStats {
  int* nsteps;  // per cell (always)
  int* nrhs;    // optional per-cell counters (may be null)
  int* njac;
  int* nlu;
  int* nnewton;
  int* nreject;
};

}  // namespace ppc
