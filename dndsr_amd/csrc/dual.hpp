// dual.hpp -- forward-mode dual numbers for exact flux derivatives.
//
// A Dual carries a value and one directional derivative.  The Riemann solvers in
// gas.hpp are templated on their scalar type, so evaluating them on Duals seeded
// with a unit direction yields the corresponding column of the flux Jacobian
// exactly (to rounding), including every entropy-fix branch and min / max switch
// (derivatives of a piecewise expression are those of the active piece).
#pragma once
#include "dev_common.hpp"

namespace dndsr
{
// the double overloads stay visible next to the Dual ones declared below
using ::fabs;
using ::fmax;
using ::fmin;
using ::sqrt;

struct Dual
{
    real v, d;
    __host__ __device__ Dual() : v(0), d(0) {}
    __host__ __device__ Dual(real x) : v(x), d(0) {}
    __host__ __device__ Dual(real x, real dx) : v(x), d(dx) {}
    __host__ __device__ Dual &operator+=(const Dual &b) { return v += b.v, d += b.d, *this; }
    __host__ __device__ Dual &operator-=(const Dual &b) { return v -= b.v, d -= b.d, *this; }
    __host__ __device__ Dual &operator*=(const Dual &b) { return d = d * b.v + v * b.d, v *= b.v, *this; }
    __host__ __device__ Dual &operator/=(const Dual &b)
    {
        const real r = 1.0 / b.v;
        d = (d - v * r * b.d) * r;
        v *= r;
        return *this;
    }
};
__host__ __device__ inline Dual operator-(const Dual &a) { return {-a.v, -a.d}; }
__host__ __device__ inline Dual operator+(const Dual &a, const Dual &b) { return {a.v + b.v, a.d + b.d}; }
__host__ __device__ inline Dual operator+(const Dual &a, real b) { return {a.v + b, a.d}; }
__host__ __device__ inline Dual operator+(real a, const Dual &b) { return {a + b.v, b.d}; }
__host__ __device__ inline Dual operator-(const Dual &a, const Dual &b) { return {a.v - b.v, a.d - b.d}; }
__host__ __device__ inline Dual operator-(const Dual &a, real b) { return {a.v - b, a.d}; }
__host__ __device__ inline Dual operator-(real a, const Dual &b) { return {a - b.v, -b.d}; }
__host__ __device__ inline Dual operator*(const Dual &a, const Dual &b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__host__ __device__ inline Dual operator*(const Dual &a, real b) { return {a.v * b, a.d * b}; }
__host__ __device__ inline Dual operator*(real a, const Dual &b) { return {a * b.v, a * b.d}; }
__host__ __device__ inline Dual operator/(const Dual &a, const Dual &b)
{
    const real r = 1.0 / b.v, q = a.v * r;
    return {q, (a.d - q * b.d) * r};
}
__host__ __device__ inline Dual operator/(const Dual &a, real b) { return {a.v / b, a.d / b}; }
__host__ __device__ inline Dual operator/(real a, const Dual &b)
{
    const real r = 1.0 / b.v, q = a * r;
    return {q, -q * b.d * r};
}
#define DNDSR_DUAL_CMP(OP)                                                                                    \
    __host__ __device__ inline bool operator OP(const Dual &a, const Dual &b) { return a.v OP b.v; }         \
    __host__ __device__ inline bool operator OP(const Dual &a, real b) { return a.v OP b; }                  \
    __host__ __device__ inline bool operator OP(real a, const Dual &b) { return a OP b.v; }
DNDSR_DUAL_CMP(<)
DNDSR_DUAL_CMP(>)
DNDSR_DUAL_CMP(<=)
DNDSR_DUAL_CMP(>=)
#undef DNDSR_DUAL_CMP
__host__ __device__ inline Dual sqrt(const Dual &a)
{
    const real s = ::sqrt(a.v);
    return {s, a.d * (0.5 / s)};
}
// a^k for a > 0
__host__ __device__ inline Dual pow(const Dual &a, real k)
{
    const real p = ::pow(a.v, k);
    return {p, a.d * (k * p / a.v)};
}
__host__ __device__ inline Dual fabs(const Dual &a) { return a.v < 0 ? -a : a; }
__host__ __device__ inline Dual fmax(const Dual &a, const Dual &b) { return a.v >= b.v ? a : b; }
__host__ __device__ inline Dual fmin(const Dual &a, const Dual &b) { return a.v <= b.v ? a : b; }
__host__ __device__ inline Dual fmax(const Dual &a, real b) { return fmax(a, Dual(b)); }
__host__ __device__ inline Dual fmax(real a, const Dual &b) { return fmax(Dual(a), b); }
__host__ __device__ inline Dual fmin(const Dual &a, real b) { return fmin(a, Dual(b)); }
__host__ __device__ inline Dual fmin(real a, const Dual &b) { return fmin(Dual(a), b); }

} // namespace dndsr
