// dual.h -- forward-mode dual numbers for the device physics templates.
//
// Dual<N> carries a value and N directional derivatives.  The pointwise
// routines of physics.h (Riemann solvers, equation of state, limiters) are
// templates on their arithmetic type, so instantiating them with Dual<N>
// evaluates the routine and N columns of its Jacobian in one pass, with the
// same branch decisions as the plain evaluation (branches compare values).
// The adjoint and tangent-linear kernels use this for the local Jacobians of
// the face fluxes; nothing in the forward solver does.
#pragma once

#include <hip/hip_runtime.h>

namespace amun {

template <int N>
struct Dual {
  double v;
  double d[N];
  __device__ __forceinline__ Dual() : v(0.0) {
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = 0.0;
  }
  __device__ __forceinline__ Dual(double x) : v(x) {
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = 0.0;
  }
  __device__ __forceinline__ Dual& operator+=(const Dual& b) {
    v += b.v;
#pragma unroll
    for (int i = 0; i < N; i++) d[i] += b.d[i];
    return *this;
  }
  __device__ __forceinline__ Dual& operator-=(const Dual& b) {
    v -= b.v;
#pragma unroll
    for (int i = 0; i < N; i++) d[i] -= b.d[i];
    return *this;
  }
  __device__ __forceinline__ Dual& operator*=(const Dual& b) {
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = d[i] * b.v + v * b.d[i];
    v *= b.v;
    return *this;
  }
};

template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  const double ib = 1.0 / b.v;
  r.v = a.v * ib;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
// mixed operands (double constants of the physics code)
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N>
__device__ __forceinline__ Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N>
__device__ __forceinline__ Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, double b) {
  Dual<N> r;
  r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(double a, const Dual<N>& b) { return b * a; }
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, double b) { return a * (1.0 / b); }

template <int N>
__device__ __forceinline__ bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N>
__device__ __forceinline__ bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
template <int N>
__device__ __forceinline__ bool operator<=(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v; }
template <int N>
__device__ __forceinline__ bool operator>=(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v; }
template <int N>
__device__ __forceinline__ bool operator==(const Dual<N>& a, const Dual<N>& b) { return a.v == b.v; }
template <int N>
__device__ __forceinline__ bool operator!=(const Dual<N>& a, const Dual<N>& b) { return a.v != b.v; }

// the helpers the physics templates call
template <int N>
__device__ __forceinline__ Dual<N> tsqrt(Dual<N> x) {
  Dual<N> r;
  r.v = sqrt(x.v);
  const double h = 0.5 / r.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = x.d[i] * h;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> rcp(Dual<N> x) {
  Dual<N> r;
  r.v = 1.0 / x.v;
  const double m = -r.v * r.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = x.d[i] * m;
  return r;
}
// copysign(x, a) = |x| sgn(a): derivative sgn(x) sgn(a) x'
template <int N>
__device__ __forceinline__ Dual<N> tsign(Dual<N> x, Dual<N> a) {
  Dual<N> r;
  r.v = copysign(x.v, a.v);
  const double s = (signbit(x.v) != signbit(a.v)) ? -1.0 : 1.0;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = s * x.d[i];
  return r;
}

}  // namespace amun
