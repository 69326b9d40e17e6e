// Forward-mode dual numbers carrying N tangent directions, for the tangent
// (JVP) and batched colored-seed kernels of the DNC5 residual.
//
// Non-smooth points follow the conventions of the TAPENADE tangent that the
// reference ships (srcfv/tangent/flux_num_dnc5_d.f90), so branch decisions and
// derivative values agree with it:
//   abs(x)   : x >= 0 takes the + branch                      (:945-951)
//   sqrt(x)  : derivative 0 where x == 0                       (:953-957)
//   max(a,b) : a < b ? b : a                                   (:1054-1074)
//   tanh     : 1 - tanh^2                                      (:1025-1026)
#pragma once
#include <hip/hip_runtime.h>

namespace bca {

template <int N>
struct Dual {
  double v;
  double d[N];
  __host__ __device__ Dual() : v(0.0) {
#pragma unroll
    for (int k = 0; k < N; ++k) d[k] = 0.0;
  }
  __host__ __device__ Dual(double a) : v(a) {
#pragma unroll
    for (int k = 0; k < N; ++k) d[k] = 0.0;
  }
};

__host__ __device__ inline double val(double a) { return a; }
template <int N> __host__ __device__ inline double val(const Dual<N>& a) { return a.v; }

#define BCA_D2(op_expr_v, op_expr_d)                                   \
  Dual<N> r;                                                           \
  r.v = (op_expr_v);                                                   \
  _Pragma("unroll") for (int k = 0; k < N; ++k) r.d[k] = (op_expr_d); \
  return r;

template <int N> __host__ __device__ inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) { BCA_D2(a.v + b.v, a.d[k] + b.d[k]) }
template <int N> __host__ __device__ inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) { BCA_D2(a.v - b.v, a.d[k] - b.d[k]) }
template <int N> __host__ __device__ inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) { BCA_D2(a.v * b.v, a.d[k] * b.v + a.v * b.d[k]) }
template <int N> __host__ __device__ inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  const double q = a.v / b.v;
  const double ib = 1.0 / b.v;
  BCA_D2(q, (a.d[k] - q * b.d[k]) * ib)
}
template <int N> __host__ __device__ inline Dual<N> operator-(const Dual<N>& a) { BCA_D2(-a.v, -a.d[k]) }
template <int N> __host__ __device__ inline Dual<N> operator+(double a, const Dual<N>& b) { BCA_D2(a + b.v, b.d[k]) }
template <int N> __host__ __device__ inline Dual<N> operator+(const Dual<N>& a, double b) { BCA_D2(a.v + b, a.d[k]) }
template <int N> __host__ __device__ inline Dual<N> operator-(double a, const Dual<N>& b) { BCA_D2(a - b.v, -b.d[k]) }
template <int N> __host__ __device__ inline Dual<N> operator-(const Dual<N>& a, double b) { BCA_D2(a.v - b, a.d[k]) }
template <int N> __host__ __device__ inline Dual<N> operator*(double a, const Dual<N>& b) { BCA_D2(a * b.v, a * b.d[k]) }
template <int N> __host__ __device__ inline Dual<N> operator*(const Dual<N>& a, double b) { BCA_D2(a.v * b, a.d[k] * b) }
template <int N> __host__ __device__ inline Dual<N> operator/(const Dual<N>& a, double b) {
  const double ib = 1.0 / b;
  BCA_D2(a.v / b, a.d[k] * ib)
}
template <int N> __host__ __device__ inline Dual<N> operator/(double a, const Dual<N>& b) {
  const double q = a / b.v;
  const double ib = 1.0 / b.v;
  BCA_D2(q, -q * b.d[k] * ib)
}

__device__ inline double dsqrt(double a) { return sqrt(a); }
template <int N> __device__ inline Dual<N> dsqrt(const Dual<N>& a) {
  const double s = sqrt(a.v);
  const double g = (a.v == 0.0) ? 0.0 : 0.5 / s;
  BCA_D2(s, a.d[k] * g)
}
__device__ inline double dtanh(double a) { return tanh(a); }
template <int N> __device__ inline Dual<N> dtanh(const Dual<N>& a) {
  const double t = tanh(a.v);
  const double g = 1.0 - t * t;
  BCA_D2(t, a.d[k] * g)
}
// x ** e with a real constant exponent
__device__ inline double dpow(double a, double e) { return pow(a, e); }
template <int N> __device__ inline Dual<N> dpow(const Dual<N>& a, double e) {
  const double p = pow(a.v, e);
  const double g = (a.v == 0.0) ? 0.0 : e * pow(a.v, e - 1.0);
  BCA_D2(p, a.d[k] * g)
}
template <class T> __device__ inline T dabs(const T& a) { return val(a) >= 0.0 ? a : -a; }
template <class T> __device__ inline T dmax(const T& a, const T& b) { return val(a) < val(b) ? b : a; }

#undef BCA_D2

}  // namespace bca
