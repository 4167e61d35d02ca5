/*
 * mgf_dual.h -- the scalar type of the FP64 path (internal).
 *
 * real_t is double in the solver library.  The sensitivity library builds
 * the same FP64 kernels and driver with real_t = mgf_dual, a dual number
 * (v, d): every operation carries the derivative of its result with respect
 * to one seeded input direction alongside the value (forward-mode
 * differentiation of the discrete iteration).  The value part is computed
 * with exactly the double operation, so values are the FP64 path's bits.
 * Derivative rules: sum, difference, product and quotient rules; fabs and
 * fmax follow the branch the value takes (at a tie fmax keeps its first
 * argument, fabs(0) keeps the derivative's sign as is).
 */
#ifndef MGF_DUAL_H
#define MGF_DUAL_H

#include <hip/hip_runtime.h>
#include <math.h>

struct mgf_dual {
    double v, d;
    mgf_dual() = default;
    __host__ __device__ constexpr mgf_dual(double x) : v(x), d(0.0) {}
    __host__ __device__ constexpr mgf_dual(double x, double y) : v(x), d(y) {}
};

#define MGF_DI __host__ __device__ __forceinline__

MGF_DI mgf_dual operator+(mgf_dual a, mgf_dual b) { return mgf_dual(a.v + b.v, a.d + b.d); }
MGF_DI mgf_dual operator-(mgf_dual a, mgf_dual b) { return mgf_dual(a.v - b.v, a.d - b.d); }
MGF_DI mgf_dual operator-(mgf_dual a) { return mgf_dual(-a.v, -a.d); }
MGF_DI mgf_dual operator*(mgf_dual a, mgf_dual b)
{
    return mgf_dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
MGF_DI mgf_dual operator/(mgf_dual a, mgf_dual b)
{
    const double q = a.v / b.v;
    return mgf_dual(q, (a.d - q * b.d) / b.v);
}
MGF_DI mgf_dual operator+(mgf_dual a, double b) { return mgf_dual(a.v + b, a.d); }
MGF_DI mgf_dual operator+(double a, mgf_dual b) { return mgf_dual(a + b.v, b.d); }
MGF_DI mgf_dual operator-(mgf_dual a, double b) { return mgf_dual(a.v - b, a.d); }
MGF_DI mgf_dual operator-(double a, mgf_dual b) { return mgf_dual(a - b.v, -b.d); }
MGF_DI mgf_dual operator*(mgf_dual a, double b) { return mgf_dual(a.v * b, a.d * b); }
MGF_DI mgf_dual operator*(double a, mgf_dual b) { return mgf_dual(a * b.v, a * b.d); }
MGF_DI mgf_dual operator/(mgf_dual a, double b) { return mgf_dual(a.v / b, a.d / b); }
MGF_DI mgf_dual operator/(double a, mgf_dual b)
{
    const double q = a / b.v;
    return mgf_dual(q, -(q * b.d) / b.v);
}
MGF_DI mgf_dual &operator+=(mgf_dual &a, mgf_dual b) { return a = a + b; }
MGF_DI mgf_dual &operator-=(mgf_dual &a, mgf_dual b) { return a = a - b; }
MGF_DI bool operator<(mgf_dual a, mgf_dual b) { return a.v < b.v; }
MGF_DI bool operator>(mgf_dual a, mgf_dual b) { return a.v > b.v; }
MGF_DI bool operator<=(mgf_dual a, mgf_dual b) { return a.v <= b.v; }
MGF_DI bool operator>=(mgf_dual a, mgf_dual b) { return a.v >= b.v; }
MGF_DI bool operator==(mgf_dual a, mgf_dual b) { return a.v == b.v; }
MGF_DI bool operator!=(mgf_dual a, mgf_dual b) { return a.v != b.v; }
MGF_DI bool operator<(mgf_dual a, double b) { return a.v < b; }
MGF_DI bool operator>(mgf_dual a, double b) { return a.v > b; }
MGF_DI bool operator==(mgf_dual a, double b) { return a.v == b; }
MGF_DI mgf_dual fabs(mgf_dual a) { return a.v < 0.0 ? -a : a; }
MGF_DI mgf_dual fmax(mgf_dual a, mgf_dual b) { return (a.v >= b.v || b.v != b.v) ? a : b; }
MGF_DI double val(mgf_dual a) { return a.v; }
MGF_DI double val(double a) { return a; }
MGF_DI double der(mgf_dual a) { return a.d; }
MGF_DI double der(double) { return 0.0; }

#undef MGF_DI

#ifdef MGF_TANGENT
typedef mgf_dual real_t;
#else
typedef double real_t;
#endif

#endif
