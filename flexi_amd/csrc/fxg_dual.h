// Forward-mode dual number (value + one directional derivative) for the adjoint kernels.
//
// The point-wise Riemann flux is evaluated with this type to obtain its directional derivatives
// with respect to the two face states; the adjoint then contracts them with the flux adjoint.
// Only the operations the templated physics of fxg_physics.h uses are provided. Comparisons and
// max/abs select by value, so each evaluation differentiates the branch the forward flux takes.
#pragma once
#include <hip/hip_runtime.h>

namespace fxg {
// A namespace of its own: the overloads below are then found by argument-dependent lookup for
// Dual arguments only, and do not hide ::sqrt/::fabs/::fmax for doubles inside namespace fxg.
namespace ad {

struct Dual {
  double v, d;
  __host__ __device__ Dual() : v(0.0), d(0.0) {}
  __host__ __device__ Dual(double x) : v(x), d(0.0) {}  // constants: zero derivative
  __host__ __device__ Dual(double x, double dx) : v(x), d(dx) {}
};

__host__ __device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
__host__ __device__ inline Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }
__host__ __device__ inline Dual& operator*=(Dual& a, Dual b) { a = a * b; return a; }
__host__ __device__ inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
__host__ __device__ inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
__host__ __device__ inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
__host__ __device__ inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
__host__ __device__ inline Dual sqrt(Dual a) {
  const double s = ::sqrt(a.v);
  return Dual(s, a.d / (2.0 * s));
}
__host__ __device__ inline Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }
// ties take the first argument, as Fortran's MAX returns one of equal arguments
__host__ __device__ inline Dual fmax(Dual a, Dual b) { return a.v >= b.v ? a : b; }

}  // namespace ad
using ad::Dual;
}  // namespace fxg
