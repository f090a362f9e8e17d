// Device-side building blocks shared by the hdgx kernels: forward-mode dual
// numbers for the pointwise physics Jacobians, the Euler / Navier-Stokes
// physics plugins, and wavefront helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HDGX_DEV __device__ __forceinline__

namespace hdgx {

// 16x16x4 MFMA operands and accumulators: A is row l%16, k l/16 and B is
// k l/16, column l%16 (lane l); an FP64 accumulator's register i holds row
// 4i + l/16, column l%16.
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

HDGX_DEV d4 mfma64(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// ------------------------------------------------------------- dual numbers
template <int N>
struct Dual {
  double v;
  double d[N];
  HDGX_DEV Dual() {}
  HDGX_DEV Dual(double x) : v(x) {
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = 0.0;
  }
  HDGX_DEV static Dual seed(double x, int k) {
    Dual r(x);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (i == k) ? 1.0 : 0.0;
    return r;
  }
};
template <int N> HDGX_DEV Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> HDGX_DEV Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> HDGX_DEV Dual<N> operator-(const Dual<N>& a) { Dual<N> r; r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
template <int N> HDGX_DEV Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = fma(a.d[i], b.v, a.v * b.d[i]); return r; }
template <int N> HDGX_DEV Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; double ib = 1.0 / b.v; r.v = a.v * ib;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib; return r; }
template <int N> HDGX_DEV Dual<N> operator+(const Dual<N>& a, double b) { Dual<N> r = a; r.v += b; return r; }
template <int N> HDGX_DEV Dual<N> operator+(double b, const Dual<N>& a) { Dual<N> r = a; r.v += b; return r; }
template <int N> HDGX_DEV Dual<N> operator-(const Dual<N>& a, double b) { Dual<N> r = a; r.v -= b; return r; }
template <int N> HDGX_DEV Dual<N> operator-(double b, const Dual<N>& a) { Dual<N> r = -a; r.v += b; return r; }
template <int N> HDGX_DEV Dual<N> operator*(const Dual<N>& a, double b) { Dual<N> r; r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b; return r; }
template <int N> HDGX_DEV Dual<N> operator*(double b, const Dual<N>& a) { return a * b; }
template <int N> HDGX_DEV Dual<N> operator/(const Dual<N>& a, double b) { return a * (1.0 / b); }
template <int N> HDGX_DEV Dual<N> operator/(double a, const Dual<N>& b) { Dual<N> r; double ib = 1.0 / b.v; r.v = a * ib; double s = -r.v * ib;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = b.d[i] * s; return r; }
template <int N> HDGX_DEV Dual<N> dsqrt(const Dual<N>& a) { Dual<N> r; r.v = sqrt(a.v); double h = 0.5 / r.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * h; return r; }
// |x|; derivative sign(x) with sign(0) = 0
template <int N> HDGX_DEV Dual<N> dabs(const Dual<N>& a) { double s = (a.v > 0) ? 1.0 : ((a.v < 0) ? -1.0 : 0.0); Dual<N> r; r.v = fabs(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i]; return r; }
HDGX_DEV double dsqrt(double a) { return sqrt(a); }
HDGX_DEV double dabs(double a) { return fabs(a); }
template <int N> HDGX_DEV double val(const Dual<N>& a) { return a.v; }
HDGX_DEV double val(double a) { return a; }
// euler/getan.m:97-103 with eps > 0: l <- 0.5 (l^2/(eps c) + eps c), and its
// derivative lines written with that new l where the old one stands:
// l' <- l l'/(eps c) + 0.5 (1 - l^2/(eps c)^2) eps c'
HDGX_DEV double llf_smooth(double l, double c, double eps) { const double ec = eps * c; return 0.5 * (l * l / ec + ec); }
template <int N> HDGX_DEV Dual<N> llf_smooth(const Dual<N>& l, const Dual<N>& c, double eps) {
  const double ec = eps * c.v;
  Dual<N> r;
  r.v = 0.5 * (l.v * l.v / ec + ec);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = r.v * l.d[i] / ec + 0.5 * (1.0 - r.v * r.v / (ec * ec)) * (eps * c.d[i]);
  return r;
}

// type of an expression in operands of types A and B (double or one Dual type):
// double only when both are
template <class A, class B> struct Prom { typedef A type; };
template <class B> struct Prom<double, B> { typedef B type; };
template <class A, class B> using prom_t = typename Prom<A, B>::type;

// ------------------------------------------------------------------ physics
enum { PHYS_EULER = 0, PHYS_NS = 1 };

// Euler inviscid flux f[c][d], state (r, r*vel[nd], rE)
template <int ND, class T>
HDGX_DEV void euler_flux(double gam, const T* u, T f[ND + 2][ND]) {
  const double gam1 = gam - 1.0;
  T r1 = 1.0 / u[0];
  T vel[ND];
  T af(0.0);
#pragma unroll
  for (int i = 0; i < ND; ++i) { vel[i] = u[1 + i] * r1; af = af + 0.5 * (vel[i] * vel[i]); }
  T p = gam1 * (u[ND + 1] - u[0] * af);
  T h = (u[ND + 1] + p) * r1;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    f[0][d] = u[1 + d];
#pragma unroll
    for (int i = 0; i < ND; ++i) f[1 + i][d] = (i == d) ? u[1 + i] * vel[d] + p : u[1 + i] * vel[d];
    f[ND + 1][d] = u[1 + d] * h;
  }
}

// Navier-Stokes 2D flux; q = (rx,rux,rvx,rEx, ry,ruy,rvy,rEy) is the (minus)
// gradient of the conservative state, f = f_inv(u) + viscous terms in (u, q).
template <class TU, class TQ, class TF>
HDGX_DEV void ns2d_flux(double gam, double Re, double Pr, const TU* u, const TQ* q, TF f[4][2]) {
  const double gam1 = gam - 1.0, Re1 = 1.0 / Re, c23 = 2.0 / 3.0, fc = gam / (gam1 * Re * Pr);
  TU r1 = 1.0 / u[0];
  TU uu = u[1] * r1, vv = u[2] * r1;
  TU q2 = 0.5 * (uu * uu + vv * vv);
  TU p = gam1 * (u[3] - u[0] * q2);
  TU h = (u[3] + p) * r1;
  TF ux = (q[1] - q[0] * uu) * r1, vx = (q[2] - q[0] * vv) * r1;
  TF uy = (q[5] - q[4] * uu) * r1, vy = (q[6] - q[4] * vv) * r1;
  TF qx = uu * ux + vv * vx, qy = uu * uy + vv * vy;
  TF px = gam1 * (q[3] - q[0] * q2 - u[0] * qx), py = gam1 * (q[7] - q[4] * q2 - u[0] * qy);
  TF Tx = (px * u[0] - p * q[0]) * (r1 * r1), Ty = (py * u[0] - p * q[4]) * (r1 * r1);
  TF txx = (Re1 * c23) * (2.0 * ux - vy), txy = Re1 * (uy + vx), tyy = (Re1 * c23) * (2.0 * vy - ux);
  f[0][0] = TF(u[1]);
  f[0][1] = TF(u[2]);
  f[1][0] = u[1] * uu + p + txx;
  f[1][1] = u[1] * vv + txy;
  f[2][0] = u[2] * uu + txy;
  f[2][1] = u[2] * vv + p + tyy;
  f[3][0] = u[1] * h + uu * txx + vv * txy + fc * Tx;
  f[3][1] = u[2] * h + uu * txy + vv * tyy + fc * Ty;
}

// Roe-form normal Jacobian applied to a vector, y = A_n(uh) x, without forming
// the matrix: A_n = rl3*I + a*cc1^T + e*cc2^T with a = (1, vel, h),
// e = (0, n, un) (ns/getan.m, fbou_euler3d.c); absflag 1 gives |A_n| with the
// entropy fix of parameter eps.
template <int ND, class T, class TX, class TY>
HDGX_DEV void roe_apply(double gam, double eps, int absflag, const T* uh, const double* n, const TX* x, TY* y) {
  const double gam1 = gam - 1.0;
  constexpr int NC = ND + 2;
  T r1 = 1.0 / uh[0];
  T vel[ND];
  T af(0.0), un(0.0);
#pragma unroll
  for (int i = 0; i < ND; ++i) { vel[i] = uh[1 + i] * r1; af = af + 0.5 * (vel[i] * vel[i]); un = un + vel[i] * n[i]; }
  T p = gam1 * (uh[ND + 1] - uh[0] * af);
  T h = (uh[ND + 1] + p) * r1;
  T c2 = gam * p * r1;
  T c = dsqrt(c2);
  T rl1, rl2, rl3;
  if (absflag) {
    rl1 = dabs(un + c); rl2 = dabs(un - c); rl3 = dabs(un);
    if (eps > 0) {
      T ec = eps * c;
      if (val(rl1) < val(ec)) rl1 = 0.5 * (rl1 * rl1 / ec + ec);
      if (val(rl2) < val(ec)) rl2 = 0.5 * (rl2 * rl2 / ec + ec);
      if (val(rl3) < val(ec)) rl3 = 0.5 * (rl3 * rl3 / ec + ec);
    }
  } else {
    rl1 = un + c; rl2 = un - c; rl3 = un;
  }
  T s1 = 0.5 * (rl1 + rl2), s2 = 0.5 * (rl1 - rl2);
  T sm = s1 - rl3;
  T ic = 1.0 / c, ic2 = 1.0 / c2;
  T A1 = gam1 * sm * ic2, A2 = s2 * ic, B1 = gam1 * s2 * ic;
  // d1 = cc1 . x, d2 = cc2 . x with b = (af, -vel, 1), d = (-un, n, 0):
  // cc1 = A1*b + A2*d, cc2 = B1*b + sm*d
  TY d1 = (A1 * af - A2 * un) * x[0], d2 = (B1 * af - sm * un) * x[0];
#pragma unroll
  for (int k = 1; k <= ND; ++k) {
    d1 = d1 + (A2 * n[k - 1] - A1 * vel[k - 1]) * x[k];
    d2 = d2 + (sm * n[k - 1] - B1 * vel[k - 1]) * x[k];
  }
  d1 = d1 + A1 * x[NC - 1];
  d2 = d2 + B1 * x[NC - 1];
  y[0] = rl3 * x[0] + d1;
#pragma unroll
  for (int i = 0; i < ND; ++i) y[1 + i] = rl3 * x[1 + i] + vel[i] * d1 + n[i] * d2;
  y[ND + 1] = rl3 * x[ND + 1] + h * d1 + un * d2;
}

// ----------------------------------------------------------- wave helpers
HDGX_DEV int wave_or(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off, 64);
  return v;
}

HDGX_DEV bool finite(double x) { return isfinite(x); }

}  // namespace hdgx
