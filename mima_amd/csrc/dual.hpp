// Forward-mode dual numbers for the per-layer local Jacobians of the backward
// (vector-Jacobian) kernels: Dual<F, N> carries a value and N directional
// derivatives.  The gas-optics and two-stream layer code is templated on its
// arithmetic type, so running it on Dual<F, N> yields the value (bitwise the
// same operations as the forward on F) and the derivatives with respect to N
// seeded inputs (layer temperature, water vapour, ozone; or the layer's gas and
// Rayleigh optical depths).
//
// Branches compare values only, and integer conversions truncate the value, so
// table indices and branch choices are those of the forward; derivatives are
// those of the branch taken (piecewise-smooth differentiation).
#pragma once

#include <hip/hip_runtime.h>

namespace rrtmg {

// the dual overloads below would hide the global math functions inside this
// namespace; bring those into scope beside them
using ::exp;
using ::fabs;
using ::fmax;
using ::fmin;
using ::log;
using ::pow;
using ::sqrt;

// a / b for the parts of a dual number: FP64 exactly; FP32 as a times the
// hardware reciprocal (1 ulp), the FP32 forward's division
__host__ __device__ __forceinline__ double dual_div(double a, double b) { return a / b; }
__device__ __forceinline__ float dual_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

template <typename F, int N>
struct Dual {
    F v;
    F d[N];
    __host__ __device__ Dual() = default;
    __host__ __device__ constexpr Dual(F x) : v(x), d{} {}
    __host__ __device__ explicit operator int() const { return (int)v; }
    __host__ __device__ static Dual seed(F x, int k)
    {
        Dual r(x);
        r.d[k] = (F)1;
        return r;
    }
};

template <typename T>
struct is_dual {
    static constexpr bool value = false;
};
template <typename F, int N>
struct is_dual<Dual<F, N>> {
    static constexpr bool value = true;
};

// value part of a plain or dual number
template <typename F>
__host__ __device__ __forceinline__ F val(F x)
{
    return x;
}
template <typename F, int N>
__host__ __device__ __forceinline__ F val(const Dual<F, N> &x)
{
    return x.v;
}

#define DUAL_T template <typename F, int N>
#define D Dual<F, N>

DUAL_T __host__ __device__ __forceinline__ D operator-(const D &a)
{
    D r;
    r.v = -a.v;
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
DUAL_T __host__ __device__ __forceinline__ D operator+(const D &a, const D &b)
{
    D r;
    r.v = a.v + b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
DUAL_T __host__ __device__ __forceinline__ D operator-(const D &a, const D &b)
{
    D r;
    r.v = a.v - b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
DUAL_T __host__ __device__ __forceinline__ D operator*(const D &a, const D &b)
{
    D r;
    r.v = a.v * b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
DUAL_T __host__ __device__ __forceinline__ D operator/(const D &a, const D &b)
{
    D r;
    r.v = dual_div(a.v, b.v);
    const F ib = dual_div((F)1, b.v);
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
// mixed dual / scalar
DUAL_T __host__ __device__ __forceinline__ D operator+(const D &a, F b)
{
    D r = a;
    r.v = a.v + b;
    return r;
}
DUAL_T __host__ __device__ __forceinline__ D operator+(F a, const D &b) { return b + a; }
DUAL_T __host__ __device__ __forceinline__ D operator-(const D &a, F b)
{
    D r = a;
    r.v = a.v - b;
    return r;
}
DUAL_T __host__ __device__ __forceinline__ D operator-(F a, const D &b)
{
    D r;
    r.v = a - b.v;
    for (int i = 0; i < N; ++i) r.d[i] = -b.d[i];
    return r;
}
DUAL_T __host__ __device__ __forceinline__ D operator*(const D &a, F b)
{
    D r;
    r.v = a.v * b;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
    return r;
}
DUAL_T __host__ __device__ __forceinline__ D operator*(F a, const D &b)
{
    D r;
    r.v = a * b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a * b.d[i];
    return r;
}
DUAL_T __host__ __device__ __forceinline__ D operator/(const D &a, F b)
{
    D r;
    r.v = dual_div(a.v, b);
    const F ib = dual_div((F)1, b);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * ib;
    return r;
}
DUAL_T __host__ __device__ __forceinline__ D operator/(F a, const D &b)
{
    // -(a/b) (db/b), associated so that no intermediate overflows where the
    // result does not (a/b^2 does for b = 1e-20 in FP32)
    D r;
    r.v = dual_div(a, b.v);
    const F ib = dual_div((F)1, b.v);
    for (int i = 0; i < N; ++i) r.d[i] = -r.v * (b.d[i] * ib);
    return r;
}
DUAL_T __host__ __device__ __forceinline__ D &operator+=(D &a, const D &b) { return a = a + b; }
DUAL_T __host__ __device__ __forceinline__ D &operator-=(D &a, const D &b) { return a = a - b; }
DUAL_T __host__ __device__ __forceinline__ D &operator*=(D &a, const D &b) { return a = a * b; }

// comparisons on the value
DUAL_T __host__ __device__ __forceinline__ bool operator<(const D &a, const D &b) { return a.v < b.v; }
DUAL_T __host__ __device__ __forceinline__ bool operator>(const D &a, const D &b) { return a.v > b.v; }
DUAL_T __host__ __device__ __forceinline__ bool operator<=(const D &a, const D &b) { return a.v <= b.v; }
DUAL_T __host__ __device__ __forceinline__ bool operator>=(const D &a, const D &b) { return a.v >= b.v; }
DUAL_T __host__ __device__ __forceinline__ bool operator==(const D &a, const D &b) { return a.v == b.v; }
DUAL_T __host__ __device__ __forceinline__ bool operator<(const D &a, F b) { return a.v < b; }
DUAL_T __host__ __device__ __forceinline__ bool operator>(const D &a, F b) { return a.v > b; }
DUAL_T __host__ __device__ __forceinline__ bool operator<=(const D &a, F b) { return a.v <= b; }
DUAL_T __host__ __device__ __forceinline__ bool operator>=(const D &a, F b) { return a.v >= b; }
DUAL_T __host__ __device__ __forceinline__ bool operator==(const D &a, F b) { return a.v == b; }
DUAL_T __host__ __device__ __forceinline__ bool operator<(F a, const D &b) { return a < b.v; }
DUAL_T __host__ __device__ __forceinline__ bool operator>(F a, const D &b) { return a > b.v; }

// elementary functions
DUAL_T __device__ __forceinline__ D sqrt(const D &a)
{
    D r;
    r.v = sqrt(a.v);
    const F s = r.v > (F)0 ? dual_div((F)0.5, r.v) : (F)0;
    for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i];
    return r;
}
DUAL_T __device__ __forceinline__ D exp(const D &a)
{
    D r;
    r.v = exp(a.v);
    for (int i = 0; i < N; ++i) r.d[i] = r.v * a.d[i];
    return r;
}
DUAL_T __device__ __forceinline__ D log(const D &a)
{
    D r;
    r.v = log(a.v);
    const F s = dual_div((F)1, a.v);
    for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i];
    return r;
}
DUAL_T __device__ __forceinline__ D pow(const D &a, F e)
{
    D r;
    r.v = pow(a.v, e);
    const F s = a.v != (F)0 ? dual_div(e * r.v, a.v) : (F)0;
    for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i];
    return r;
}
DUAL_T __device__ __forceinline__ D fabs(const D &a) { return a.v < (F)0 ? -a : a; }
DUAL_T __device__ __forceinline__ D fmin(const D &a, F b) { return a.v <= b ? a : D(b); }
DUAL_T __device__ __forceinline__ D fmax(const D &a, F b) { return a.v >= b ? a : D(b); }
DUAL_T __device__ __forceinline__ D fmin(const D &a, const D &b) { return a.v <= b.v ? a : b; }
DUAL_T __device__ __forceinline__ D fmax(const D &a, const D &b) { return a.v >= b.v ? a : b; }

#undef D
#undef DUAL_T

// convert between arithmetic types (value and derivatives)
template <typename To, typename From>
struct Cast {
    __host__ __device__ static To to(const From &x) { return (To)x; }
};
template <typename T, typename S, int N>
struct Cast<Dual<T, N>, Dual<S, N>> {
    __host__ __device__ static Dual<T, N> to(const Dual<S, N> &x)
    {
        Dual<T, N> r;
        r.v = (T)x.v;
        for (int i = 0; i < N; ++i) r.d[i] = (T)x.d[i];
        return r;
    }
};
template <typename To, typename From>
__host__ __device__ __forceinline__ To cast(const From &x)
{
    return Cast<To, From>::to(x);
}

}  // namespace rrtmg
