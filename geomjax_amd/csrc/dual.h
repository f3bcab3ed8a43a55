// Forward-mode dual numbers for the integrator kernels: a value and NT
// tangents.  Instantiating a kernel with Dual<double, 2> propagates the exact
// derivatives of every executed operation with respect to two seeded inputs
// (here log step size and log alpha^2), the semantics of jax.jacfwd through
// the executed iterates.  Comparisons and branch decisions use the value only.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "common.h"

namespace gjx {

// keep the plain math functions visible next to the dual overloads below
using ::exp;
using ::fabs;
using ::log;
using ::sqrt;

template <class T, int NT>
struct Dual {
  T v;
  T d[NT];
  __host__ __device__ Dual() : v(0) {
    for (int i = 0; i < NT; ++i) d[i] = 0;
  }
  __host__ __device__ Dual(T x) : v(x) {  // NOLINT: implicit lift of constants
    for (int i = 0; i < NT; ++i) d[i] = 0;
  }
  __host__ __device__ explicit Dual(double x, int) : v((T)x) {
    for (int i = 0; i < NT; ++i) d[i] = 0;
  }
};

template <class T, int N> __device__ __forceinline__ Dual<T, N> operator+(Dual<T, N> a, const Dual<T, N>& b) {
  a.v += b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) a.d[i] += b.d[i];
  return a;
}
template <class T, int N> __device__ __forceinline__ Dual<T, N> operator-(Dual<T, N> a, const Dual<T, N>& b) {
  a.v -= b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) a.d[i] -= b.d[i];
  return a;
}
template <class T, int N> __device__ __forceinline__ Dual<T, N> operator-(Dual<T, N> a) {
  a.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) a.d[i] = -a.d[i];
  return a;
}
template <class T, int N> __device__ __forceinline__ Dual<T, N> operator*(const Dual<T, N>& a, const Dual<T, N>& b) {
  Dual<T, N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <class T, int N> __device__ __forceinline__ Dual<T, N> operator/(const Dual<T, N>& a, const Dual<T, N>& b) {
  Dual<T, N> r;
  const T ib = T(1) / b.v;
  r.v = a.v * ib;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
// mixed with plain scalars
template <class T, int N> __device__ __forceinline__ Dual<T, N> operator+(Dual<T, N> a, T b) { a.v += b; return a; }
template <class T, int N> __device__ __forceinline__ Dual<T, N> operator+(T b, Dual<T, N> a) { a.v += b; return a; }
template <class T, int N> __device__ __forceinline__ Dual<T, N> operator-(Dual<T, N> a, T b) { a.v -= b; return a; }
template <class T, int N> __device__ __forceinline__ Dual<T, N> operator-(T b, const Dual<T, N>& a) { return Dual<T, N>(b) - a; }
template <class T, int N> __device__ __forceinline__ Dual<T, N> operator*(Dual<T, N> a, T b) {
  a.v *= b;
#pragma unroll
  for (int i = 0; i < N; ++i) a.d[i] *= b;
  return a;
}
template <class T, int N> __device__ __forceinline__ Dual<T, N> operator*(T b, Dual<T, N> a) { return a * b; }
template <class T, int N> __device__ __forceinline__ Dual<T, N> operator/(const Dual<T, N>& a, T b) { return a * (T(1) / b); }
template <class T, int N> __device__ __forceinline__ Dual<T, N> operator/(T a, const Dual<T, N>& b) { return Dual<T, N>(a) / b; }

template <class T, int N> __device__ __forceinline__ Dual<T, N>& operator+=(Dual<T, N>& a, const Dual<T, N>& b) { a = a + b; return a; }
template <class T, int N> __device__ __forceinline__ Dual<T, N>& operator-=(Dual<T, N>& a, const Dual<T, N>& b) { a = a - b; return a; }
template <class T, int N> __device__ __forceinline__ Dual<T, N>& operator*=(Dual<T, N>& a, const Dual<T, N>& b) { a = a * b; return a; }

template <class T, int N> __device__ __forceinline__ bool operator<(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v < b.v; }
template <class T, int N> __device__ __forceinline__ bool operator>(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v > b.v; }
template <class T, int N> __device__ __forceinline__ bool operator!=(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v != b.v; }

template <class T, int N> __device__ __forceinline__ Dual<T, N> exp(const Dual<T, N>& a) {
  Dual<T, N> r;
  r.v = exp(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = r.v * a.d[i];
  return r;
}
template <class T, int N> __device__ __forceinline__ Dual<T, N> log(const Dual<T, N>& a) {
  Dual<T, N> r;
  r.v = log(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / a.v;
  return r;
}
template <class T, int N> __device__ __forceinline__ Dual<T, N> sqrt(const Dual<T, N>& a) {
  Dual<T, N> r;
  r.v = sqrt(a.v);
  const T h = T(0.5) / r.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * h;
  return r;
}
template <class T, int N> __device__ __forceinline__ Dual<T, N> fabs(const Dual<T, N>& a) { return a.v < T(0) ? -a : a; }

// value of a plain or dual number
__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ float val(float x) { return x; }
template <class T, int N> __device__ __forceinline__ T val(const Dual<T, N>& x) { return x.v; }

// group reductions of dual numbers: component-wise sums
template <int W, class T, int N>
__device__ __forceinline__ Dual<T, N> group_sum(Dual<T, N> x) {
  x.v = group_sum<W>(x.v);
#pragma unroll
  for (int i = 0; i < N; ++i) x.d[i] = group_sum<W>(x.d[i]);
  return x;
}

template <int W, class T, int N>
__device__ __forceinline__ Dual<T, N> group_exclusive_scan(Dual<T, N> x, int lane_in_group) {
  x.v = group_exclusive_scan<W>(x.v, lane_in_group);
#pragma unroll
  for (int i = 0; i < N; ++i) x.d[i] = group_exclusive_scan<W>(x.d[i], lane_in_group);
  return x;
}

// Scalar type of float64 energies for a state type R: double for plain
// kernels, the dual itself for tangent kernels.
template <class R> struct EnergyOf { using type = double; };
template <class T, int N> struct EnergyOf<Dual<T, N>> { using type = Dual<double, N>; };

// widen a state scalar to the energy type
__device__ __forceinline__ double widen(double x) { return x; }
__device__ __forceinline__ double widen(float x) { return (double)x; }
template <int N> __device__ __forceinline__ Dual<double, N> widen(const Dual<double, N>& x) { return x; }
template <int N> __device__ __forceinline__ Dual<double, N> widen(const Dual<float, N>& x) {
  Dual<double, N> r((double)x.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (double)x.d[i];
  return r;
}

}  // namespace gjx
