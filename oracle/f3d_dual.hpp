// Forward-mode dual numbers for the CPU oracle (test infrastructure): a + b*eps with eps^2 = 0.
// Instantiating the oracle's templates with Dual gives the Jacobian-vector product of the
// reference's RHS (the CPU forward-mode reference gradient).  Comparisons and branch decisions
// use the value only, so the scheme mask, minmod branches and tangent choice are frozen at the
// primal point, as in the adjoint.
#pragma once
#include <cmath>

namespace f3d_oracle {

struct Dual {
  double v, d;
  Dual() : v(0), d(0) {}
  Dual(double x) : v(x), d(0) {}
  Dual(double x, double y) : v(x), d(y) {}
  explicit operator double() const { return v; }
  explicit operator float() const { return (float)v; }
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
inline Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
inline Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
#define F3D_DUAL_MIXED(op)                                         \
  inline Dual operator op(Dual a, double b) { return a op Dual(b); } \
  inline Dual operator op(double a, Dual b) { return Dual(a) op b; } \
  inline Dual operator op(Dual a, int b) { return a op Dual((double)b); } \
  inline Dual operator op(int a, Dual b) { return Dual((double)a) op b; }
F3D_DUAL_MIXED(+)
F3D_DUAL_MIXED(-)
F3D_DUAL_MIXED(*)
F3D_DUAL_MIXED(/)
#undef F3D_DUAL_MIXED
#define F3D_DUAL_CMP(op)                                                  \
  inline bool operator op(Dual a, Dual b) { return a.v op b.v; }           \
  inline bool operator op(Dual a, double b) { return a.v op b; }           \
  inline bool operator op(double a, Dual b) { return a op b.v; }           \
  inline bool operator op(Dual a, int b) { return a.v op (double)b; }
F3D_DUAL_CMP(<)
F3D_DUAL_CMP(>)
F3D_DUAL_CMP(<=)
F3D_DUAL_CMP(>=)
F3D_DUAL_CMP(==)
F3D_DUAL_CMP(!=)
#undef F3D_DUAL_CMP
// |x|: derivative sign(x) with sign(0) = 0 (the subgradient the adjoint uses at x = 0)
inline Dual fabs(Dual a) { return {std::fabs(a.v), a.v > 0 ? a.d : (a.v < 0 ? -a.d : 0.0)}; }
inline Dual sqrt(Dual a) {
  const double s = std::sqrt(a.v);
  return {s, a.d / (2 * s)};
}

}  // namespace f3d_oracle

namespace std {
inline f3d_oracle::Dual fabs(f3d_oracle::Dual a) { return f3d_oracle::fabs(a); }
inline f3d_oracle::Dual sqrt(f3d_oracle::Dual a) { return f3d_oracle::sqrt(a); }
template <>
struct numeric_limits<f3d_oracle::Dual> : numeric_limits<double> {};
}  // namespace std
