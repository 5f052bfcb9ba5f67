// dual.h -- forward-mode dual numbers for the pointwise functional.
//
// phi_grad<dual, dual, NC, ND> evaluated at n + eps u returns, in the eps parts of
// its gradient, the Hessian-vector product H_Phi(n) u: the second-derivative
// kernel behind the adjoint of the fixed point.  Only the operations phi_grad,
// assoc_grad and i_lite use are provided; comparisons and selections act on the
// value part.
#pragma once
#include <hip/hip_runtime.h>

namespace cdft {

// keep the floating-point overloads visible next to the dual ones below
using ::copysign;
using ::fabs;
using ::fmax;
using ::log;
using ::sqrt;

struct dual {
    double v, d;
    __host__ __device__ constexpr dual() : v(0.0), d(0.0) {}
    __host__ __device__ constexpr dual(double x) : v(x), d(0.0) {}
    __host__ __device__ constexpr dual(double x, double dx) : v(x), d(dx) {}
};

#define CDFT_DI __host__ __device__ __attribute__((always_inline)) inline
CDFT_DI dual operator+(dual a, dual b) { return {a.v + b.v, a.d + b.d}; }
CDFT_DI dual operator-(dual a, dual b) { return {a.v - b.v, a.d - b.d}; }
CDFT_DI dual operator-(dual a) { return {-a.v, -a.d}; }
CDFT_DI dual operator*(dual a, dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
CDFT_DI dual operator/(dual a, dual b)
{
    const double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
CDFT_DI dual &operator+=(dual &a, dual b) { return a = a + b; }
CDFT_DI dual &operator-=(dual &a, dual b) { return a = a - b; }
CDFT_DI dual &operator*=(dual &a, dual b) { return a = a * b; }
CDFT_DI dual &operator/=(dual &a, dual b) { return a = a / b; }
CDFT_DI bool operator<(dual a, dual b) { return a.v < b.v; }
CDFT_DI bool operator>(dual a, dual b) { return a.v > b.v; }
CDFT_DI bool operator<=(dual a, dual b) { return a.v <= b.v; }
CDFT_DI bool operator>=(dual a, dual b) { return a.v >= b.v; }
CDFT_DI bool operator==(dual a, dual b) { return a.v == b.v; }
CDFT_DI dual log(dual a) { return {::log(a.v), a.d / a.v}; }
CDFT_DI dual sqrt(dual a)
{
    const double r = ::sqrt(a.v);
    return {r, 0.5 * a.d / r};
}
CDFT_DI dual fabs(dual a) { return a.v < 0.0 ? -a : a; }
CDFT_DI dual copysign(dual a, dual b) { return ((a.v < 0.0) != (b.v < 0.0)) ? -a : a; }
CDFT_DI dual fmax(dual a, dual b) { return a.v >= b.v ? a : b; }
#undef CDFT_DI

}  // namespace cdft
