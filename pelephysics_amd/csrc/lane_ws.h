// Per-lane matrices of the larger mechanisms in an HBM workspace.
//
// A lane's private (scratch) frame is limited to 128 KiB on gfx950, and the
// compiler's frame handling breaks down well before that when the frame holds
// several n x n matrices (dodecane_lu: n = 54, heptane_lu_88sk: n = 89).  For
// mechanisms with NEQ > PPC_WS_MIN_NEQ the integrators keep their Jacobians,
// LU factors and Krylov bases in a workspace the host allocates instead, laid
// out element-major across lanes: element i of an array that starts at row o
// lives at ws[(o + i) * lanes + g] for lane g, so the 64 lanes of a wave touch
// 64 consecutive words (one coalesced 512-byte access for doubles), the same
// interleaving the hardware gives private memory.  A row is `lanes` doubles; an
// array of m floats takes ceil(m / 2) rows.
#pragma once
#include <type_traits>
#include "erk_tables.h"
#include "ppc_common.h"
#include "reactors_fwd.h"

#ifndef PPC_WS_MIN_NEQ
#define PPC_WS_MIN_NEQ 32
#endif

namespace ppc {

template <class M>
constexpr bool uses_ws() {
  return M::NEQ > PPC_WS_MIN_NEQ;
}

// One lane's view of an array in the workspace: element i at p[i * s].
template <typename T>
struct LaneArr {
  T* p;
  long long s;
  PPC_DI T& operator[](long long i) const { return p[i * s]; }
};

struct LaneWS {
  double* base;     // workspace (lanes * rows doubles)
  long long lanes;  // lanes sharing it (a multiple of 64)
  long long g;      // this lane
  template <typename T>
  PPC_DI LaneArr<T> at(long long row) const {
    return {reinterpret_cast<T*>(base + row * lanes) + g, lanes};
  }
};

// Rows taken by m elements of T.
template <typename T>
constexpr long long ws_rows(long long m) {
  return sizeof(T) == 8 ? m : (m + 1) / 2;
}

// An array of T with m elements: private when WS is false (`priv` must hold
// m elements), else at `row` of the workspace.
template <bool WS, typename T>
PPC_DI auto lane_array(T* priv, const LaneWS& w, long long row) {
  if constexpr (WS)
    return w.template at<T>(row);
  else
    return priv;
}

// Workspace rows per lane of each integrator; 0 when the mechanism keeps its
// matrices in private memory.
template <class M>
constexpr long long integrator_ws_rows(int mode) {
  constexpr long long n = M::NEQ, nn = n * n;
  if (mode == MODE_ERK) return ERK_MAX_STAGES * n;  // the stages, every mechanism
  if (!uses_ws<M>()) return 0;
  switch (mode) {
    case MODE_BDF: return nn + 2 * (n + 1) * n;                   // A, Krylov basis V, Hessenberg H
    case MODE_ADAPTIVE: return 2 * nn;                            // J, LU (FP64)
    case MODE_BATCHED_MIXED: return 2 * ws_rows<float>(nn) + nn;  // LU (FP32), Jf, Jd
    case MODE_BATCHED_FP64: return nn + ws_rows<float>(nn) + nn;  // LU (FP64), Jf, Jd
    default: return 0;
  }
}

// ---- batched group kernel (group_bdf.h): G lanes per cell
template <class M, class = void>
struct has_table : std::false_type {};
template <class M>
struct has_table<M, std::void_t<decltype(M::kTable)>> : std::integral_constant<bool, M::kTable> {};

// Groups of G lanes per cell exist for the table-driven mechanisms with at most
// three equations per lane (NEQ <= 3 G) and at least half the lanes holding
// an equation (G <= 2 NEQ).
template <class M, class = void>
struct has_qss : std::false_type {};
template <class M>
struct has_qss<M, std::void_t<decltype(M::NQ)>> : std::integral_constant<bool, (M::NQ > 0)> {};

// (QSS mechanisms run one lane per cell.)
template <class M, int G>
constexpr bool group_ok() {
  return has_table<M>::value && !has_qss<M>::value && M::NEQ <= 3 * G && G <= 2 * M::NEQ;
}

// Device workspace of one group: the n x n FP64 Jacobian accumulator and its
// rows (at most FP64)
template <int n, int G>
constexpr long long group_ws_doubles_max() {
  return (long long)n * n + (long long)((n + G - 1) / G) * n * G;
}

// Workspace rows per lane (doubles, 64-lane blocks as lane_ws) of the group
// kernel with cells_per_wave cells per wave; 0 when there is no such kernel.
template <class M>
long long group_ws_rows(int cells_per_wave) {
  constexpr int n = M::NEQ;
  switch (cells_per_wave) {
    case 8: return group_ok<M, 8>() ? (8 * group_ws_doubles_max<n, 8>() + 63) / 64 : 0;
    case 4: return group_ok<M, 16>() ? (4 * group_ws_doubles_max<n, 16>() + 63) / 64 : 0;
    case 2: return group_ok<M, 32>() ? (2 * group_ws_doubles_max<n, 32>() + 63) / 64 : 0;
    case 1: return group_ok<M, 64>() ? (group_ws_doubles_max<n, 64>() + 63) / 64 : 0;
    default: return 0;
  }
}

}  // namespace ppc
