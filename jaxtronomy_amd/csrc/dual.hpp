// Forward-mode dual numbers for the gradient kernels: a value and N tangents,
// all in registers (every index is a compile-time constant after unrolling).
// Functions take the value's derivative from the closed form of each
// elementary function; branches and floors are decided on the value, as the
// reference's jnp.where / maximum are under JAX differentiation.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace lensim {
namespace gm {  // generic math: the same calls for float, double and Dual

template <typename T, int N>
struct Dual {
  T v;
  T d[N];
  __device__ __forceinline__ Dual() {}
  __device__ __forceinline__ Dual(T x) : v(x) {
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = T(0);
  }
  // the i-th seed variable of value x
  static __device__ __forceinline__ Dual seed(T x, int i) {
    Dual r(x);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (k == i) ? T(1) : T(0);
    return r;
  }
};

// value of a scalar or a dual
template <typename T>
__device__ __forceinline__ T val(T x) {
  return x;
}
template <typename T, int N>
__device__ __forceinline__ T val(const Dual<T, N> &x) {
  return x.v;
}

// f(x) with f'(x) = dfx: chain rule helper
template <typename T, int N>
__device__ __forceinline__ Dual<T, N> chain(const Dual<T, N> &x, T fx, T dfx) {
  Dual<T, N> r;
  r.v = fx;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = dfx * x.d[i];
  return r;
}

#define DUAL_T template <typename T, int N>
#define D Dual<T, N>

DUAL_T __device__ __forceinline__ D operator+(const D &a, const D &b) {
  D r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
DUAL_T __device__ __forceinline__ D operator-(const D &a, const D &b) {
  D r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
DUAL_T __device__ __forceinline__ D operator-(const D &a) {
  D r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
DUAL_T __device__ __forceinline__ D operator*(const D &a, const D &b) {
  D r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
DUAL_T __device__ __forceinline__ D operator/(const D &a, const D &b) {
  D r;
  const T ib = T(1) / b.v;
  r.v = a.v * ib;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
// mixed with plain scalars
DUAL_T __device__ __forceinline__ D operator+(const D &a, T b) {
  D r = a;
  r.v += b;
  return r;
}
DUAL_T __device__ __forceinline__ D operator+(T a, const D &b) { return b + a; }
DUAL_T __device__ __forceinline__ D operator-(const D &a, T b) {
  D r = a;
  r.v -= b;
  return r;
}
DUAL_T __device__ __forceinline__ D operator-(T a, const D &b) {
  D r = -b;
  r.v += a;
  return r;
}
DUAL_T __device__ __forceinline__ D operator*(const D &a, T b) {
  D r;
  r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
  return r;
}
DUAL_T __device__ __forceinline__ D operator*(T a, const D &b) { return b * a; }
DUAL_T __device__ __forceinline__ D operator/(const D &a, T b) { return a * (T(1) / b); }
DUAL_T __device__ __forceinline__ D operator/(T a, const D &b) {
  const T ib = T(1) / b.v;
  return chain(b, a * ib, -a * ib * ib);
}

// elementary functions on duals (values through the accurate library calls)
DUAL_T __device__ __forceinline__ D sqrt(const D &x) {
  const T s = ::sqrt(x.v);
  return chain(x, s, T(0.5) / s);
}
DUAL_T __device__ __forceinline__ D exp(const D &x) {
  const T e = ::exp(x.v);
  return chain(x, e, e);
}
DUAL_T __device__ __forceinline__ D log(const D &x) { return chain(x, T(::log(x.v)), T(1) / x.v); }
DUAL_T __device__ __forceinline__ D atan(const D &x) { return chain(x, T(::atan(x.v)), T(1) / (T(1) + x.v * x.v)); }
DUAL_T __device__ __forceinline__ D atanh(const D &x) {
  return chain(x, T(::atanh(x.v)), T(1) / (T(1) - x.v * x.v));
}
DUAL_T __device__ __forceinline__ D acosh(const D &x) {
  return chain(x, T(::acosh(x.v)), T(1) / ::sqrt(x.v * x.v - T(1)));
}
DUAL_T __device__ __forceinline__ D acos(const D &x) {
  return chain(x, T(::acos(x.v)), T(-1) / ::sqrt(T(1) - x.v * x.v));
}
DUAL_T __device__ __forceinline__ D sin(const D &x) { return chain(x, T(::sin(x.v)), T(::cos(x.v))); }
DUAL_T __device__ __forceinline__ D cos(const D &x) { return chain(x, T(::cos(x.v)), T(-::sin(x.v))); }
DUAL_T __device__ __forceinline__ D atan2(const D &y, const D &x) {
  // d atan2(y, x) = (x dy - y dx) / (x^2 + y^2)
  D r;
  r.v = ::atan2(y.v, x.v);
  const T ir = T(1) / (x.v * x.v + y.v * y.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) * ir;
  return r;
}
// x^p for a plain exponent
DUAL_T __device__ __forceinline__ D pow(const D &x, T p) {
  const T y = ::pow(x.v, p);
  return chain(x, y, p * y / x.v);
}
// a^b = exp(b ln a)
DUAL_T __device__ __forceinline__ D pow(const D &a, const D &b) { return exp(b * log(a)); }
// floors / clamps decided on the value (derivative of the selected branch)
DUAL_T __device__ __forceinline__ D fmax(const D &a, T b) { return a.v >= b ? a : D(b); }
DUAL_T __device__ __forceinline__ D fmin(const D &a, T b) { return a.v <= b ? a : D(b); }
DUAL_T __device__ __forceinline__ D fabs(const D &a) { return a.v >= T(0) ? a : -a; }

#undef D
#undef DUAL_T

// plain-scalar versions of the same names
#define GM_SCALAR(T)                                                                              \
  __device__ __forceinline__ T sqrt(T x) { return ::sqrt(x); }                                  \
  __device__ __forceinline__ T exp(T x) { return ::exp(x); }                                    \
  __device__ __forceinline__ T log(T x) { return ::log(x); }                                    \
  __device__ __forceinline__ T atan(T x) { return ::atan(x); }                                  \
  __device__ __forceinline__ T atanh(T x) { return ::atanh(x); }                                \
  __device__ __forceinline__ T acosh(T x) { return ::acosh(x); }                                \
  __device__ __forceinline__ T acos(T x) { return ::acos(x); }                                  \
  __device__ __forceinline__ T sin(T x) { return ::sin(x); }                                    \
  __device__ __forceinline__ T cos(T x) { return ::cos(x); }                                    \
  __device__ __forceinline__ T atan2(T y, T x) { return ::atan2(y, x); }                        \
  __device__ __forceinline__ T pow(T x, T p) { return ::pow(x, p); }                            \
  __device__ __forceinline__ T fmax(T a, T b) { return a >= b ? a : b; }                        \
  __device__ __forceinline__ T fmin(T a, T b) { return a <= b ? a : b; }                        \
  __device__ __forceinline__ T fabs(T a) { return a >= T(0) ? a : -a; }
GM_SCALAR(float)
GM_SCALAR(double)
#undef GM_SCALAR

}  // namespace gm
}  // namespace lensim
