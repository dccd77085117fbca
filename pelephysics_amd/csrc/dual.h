// Forward-mode dual numbers (value + one directional derivative) for the
// second-order terms of the discrete adjoint: d/du_l of the Newton-matrix
// Jacobian is obtained by evaluating the same table-driven Jacobian code with
// S = Dual, seeded along u_l.
#pragma once
#include "ppc_common.h"

namespace ppc {

// keep the floating-point overloads visible next to the Dual ones
using ::exp;
using ::log;
using ::log10;

struct Dual {
  double v, d;
  Dual() = default;  // trivial, so that Dual arrays can live in LDS
  PPC_HD Dual(double x) : v(x), d(0.0) {}
  PPC_HD Dual(double x, double dx) : v(x), d(dx) {}
  PPC_HD Dual& operator+=(const Dual& o) {
    v += o.v;
    d += o.d;
    return *this;
  }
  PPC_HD Dual& operator-=(const Dual& o) {
    v -= o.v;
    d -= o.d;
    return *this;
  }
  PPC_HD Dual& operator*=(const Dual& o) {
    d = d * o.v + v * o.d;
    v *= o.v;
    return *this;
  }
};

PPC_HD inline Dual operator+(Dual a, const Dual& b) { return a += b; }
PPC_HD inline Dual operator-(Dual a, const Dual& b) { return a -= b; }
PPC_HD inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
PPC_HD inline Dual operator*(Dual a, const Dual& b) { return a *= b; }
PPC_HD inline Dual operator/(const Dual& a, const Dual& b) {
  const double iv = 1.0 / b.v;
  const double q = a.v * iv;
  return Dual(q, (a.d - q * b.d) * iv);
}
PPC_HD inline Dual operator+(const Dual& a, double b) { return Dual(a.v + b, a.d); }
PPC_HD inline Dual operator+(double a, const Dual& b) { return Dual(a + b.v, b.d); }
PPC_HD inline Dual operator-(const Dual& a, double b) { return Dual(a.v - b, a.d); }
PPC_HD inline Dual operator-(double a, const Dual& b) { return Dual(a - b.v, -b.d); }
PPC_HD inline Dual operator*(const Dual& a, double b) { return Dual(a.v * b, a.d * b); }
PPC_HD inline Dual operator*(double a, const Dual& b) { return Dual(a * b.v, a * b.d); }
PPC_HD inline Dual operator/(const Dual& a, double b) { return Dual(a.v / b, a.d / b); }
PPC_HD inline Dual operator/(double a, const Dual& b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
PPC_HD inline bool operator<(const Dual& a, double b) { return a.v < b; }
PPC_HD inline bool operator>(const Dual& a, double b) { return a.v > b; }

PPC_HD inline Dual exp(const Dual& a) {
  const double e = ::exp(a.v);
  return Dual(e, e * a.d);
}
PPC_HD inline Dual log(const Dual& a) { return Dual(::log(a.v), a.d / a.v); }
PPC_HD inline Dual log10(const Dual& a) { return Dual(::log10(a.v), a.d / (a.v * LN10)); }

// value of a scalar of either type
PPC_HD inline double val(double x) { return x; }
PPC_HD inline double val(const Dual& x) { return x.v; }

}  // namespace ppc
