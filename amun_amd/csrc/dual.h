// dual.h -- forward-mode dual numbers for the device physics templates.
//
// Dual<N, S> carries a value and N directional derivatives in arithmetic S
// (double, or float for the adjoint of the FP32-flux mode).  The pointwise
// routines of physics.h (Riemann solvers, equation of state, limiters) are
// templates on their arithmetic type, so instantiating them with Dual<N>
// evaluates the routine and N columns of its Jacobian in one pass, with the
// same branch decisions as the plain evaluation (branches compare values).
// The adjoint and tangent-linear kernels use this for the local Jacobians of
// the face fluxes; nothing in the forward solver does.
#pragma once

#include <hip/hip_runtime.h>

namespace amun {

template <int N, typename S = double>
struct Dual {
  S v;
  S d[N];
  __device__ __forceinline__ Dual() : v(S(0)) {
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = S(0);
  }
  __device__ __forceinline__ Dual(double x) : v(S(x)) {
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = S(0);
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

template <int N, typename S>
__device__ __forceinline__ Dual<N, S> operator-(const Dual<N, S>& a) {
  Dual<N, S> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
  return r;
}
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> operator+(const Dual<N, S>& a, const Dual<N, S>& b) {
  Dual<N, S> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> operator-(const Dual<N, S>& a, const Dual<N, S>& b) {
  Dual<N, S> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> operator*(const Dual<N, S>& a, const Dual<N, S>& b) {
  Dual<N, S> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> operator/(const Dual<N, S>& a, const Dual<N, S>& b) {
  Dual<N, S> r;
  const S ib = S(1) / b.v;
  r.v = a.v * ib;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
// mixed operands (double constants of the physics code)
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> operator+(const Dual<N, S>& a, double b) { return a + Dual<N, S>(b); }
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> operator+(double a, const Dual<N, S>& b) { return Dual<N, S>(a) + b; }
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> operator-(const Dual<N, S>& a, double b) { return a - Dual<N, S>(b); }
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> operator-(double a, const Dual<N, S>& b) { return Dual<N, S>(a) - b; }
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> operator*(const Dual<N, S>& a, double b) {
  Dual<N, S> r;
  r.v = a.v * S(b);
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * S(b);
  return r;
}
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> operator*(double a, const Dual<N, S>& b) { return b * a; }
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> operator/(const Dual<N, S>& a, double b) { return a * (1.0 / b); }

template <int N, typename S>
__device__ __forceinline__ bool operator<(const Dual<N, S>& a, const Dual<N, S>& b) { return a.v < b.v; }
template <int N, typename S>
__device__ __forceinline__ bool operator>(const Dual<N, S>& a, const Dual<N, S>& b) { return a.v > b.v; }
template <int N, typename S>
__device__ __forceinline__ bool operator<=(const Dual<N, S>& a, const Dual<N, S>& b) { return a.v <= b.v; }
template <int N, typename S>
__device__ __forceinline__ bool operator>=(const Dual<N, S>& a, const Dual<N, S>& b) { return a.v >= b.v; }
template <int N, typename S>
__device__ __forceinline__ bool operator==(const Dual<N, S>& a, const Dual<N, S>& b) { return a.v == b.v; }
template <int N, typename S>
__device__ __forceinline__ bool operator!=(const Dual<N, S>& a, const Dual<N, S>& b) { return a.v != b.v; }

// the helpers the physics templates call
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> tsqrt(Dual<N, S> x) {
  Dual<N, S> r;
  r.v = sqrt(x.v);
  const S h = S(0.5) / r.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = x.d[i] * h;
  return r;
}
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> rcp(Dual<N, S> x) {
  Dual<N, S> r;
  r.v = S(1) / x.v;
  const S m = -r.v * r.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = x.d[i] * m;
  return r;
}
// copysign(x, a) = |x| sgn(a): derivative sgn(x) sgn(a) x'
template <int N, typename S>
__device__ __forceinline__ Dual<N, S> tsign(Dual<N, S> x, Dual<N, S> a) {
  Dual<N, S> r;
  r.v = copysign(x.v, a.v);
  const S s = (signbit(x.v) != signbit(a.v)) ? S(-1) : S(1);
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = s * x.d[i];
  return r;
}

}  // namespace amun
