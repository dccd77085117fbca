// Types shared by the kernels and the host launchers.
#pragma once
#include "ppc_common.h"
#include "ppc_types.h"

namespace ppc {

// integrators: 0 ReactorBDF parity, 1 ReactorRK64 parity, 2 adaptive BDF (plain,
// one lane per cell, FP64), 3 batched adaptive BDF with FP32 LU, 4 batched
// adaptive BDF with FP64 LU, 6 ReactorArkode's explicit Runge-Kutta
enum : int { MODE_BDF = 0, MODE_RK64 = 1, MODE_ADAPTIVE = 2, MODE_BATCHED_MIXED = 3, MODE_BATCHED_FP64 = 4,
             MODE_ERK = 6 };

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

struct QueueArgs {
  const int* order;  // cell visiting order (null: identity)
  int* counter;      // global work counter (zeroed before the launch)
  int ncells;
  int max_blocks;    // upper bound on the persistent grid (0: every resident block)
};

// Arguments of the discrete-adjoint kernel (wave_adjoint.h); device pointers.
struct AdjointArgs {
  double* y;          // ncells x NEQ (YCOrder): initial state in, final state out
  const double* rX;   // ncells: rho e (reactor_type 1) or rho h (2); no sources
  double* grad;       // ncells x NR x 3: d(T_final, Y_a, Y_b)/d ln A_j, reactions in mechanism order
  double* out;        // ncells x 3: T_final, Y_a, Y_b
  double* ckpt;       // ncells x (nmacro * nsubsteps) x NEQ checkpoints
  int ncells, nmacro, spec_a, spec_b;
  double dt;          // per call
};

// Box <-> flat state conversion (flatten.h); device pointers.
struct FlatArgs {
  BoxArgs b;             // rY, rYsrc, T, rE, rEsrc, FC views over the box
  double* y;             // ncells x (N+1) YCOrder
  double* ysrc;          // ncells x N (flatten only)
  double* e;             // ncells: rho e or rho h
  double* esrc;          // ncells (flatten only)
  const long long* fc;   // ncells step counts written to FC (unflatten only)
  const int* fc32;       // the same as int counts, used when fc is null
  double dt;             // unflatten: rE = e + dt * rEsrc
  int reactor_type, clean;
  int divide;            // energy per mass as e / rho (ReactorBDF.cpp:379,469) instead of e * (1 / rho)
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
