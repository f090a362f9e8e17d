// Oracle: CPU FP64 restatement of the reference's HDG elemental path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load this library, as the checker and as the
// timed CPU baseline.  The product (hdg_amd/) never links or calls it.
//
// Follows, function by function:
//   volgeom          kernel/volgeom.m:22-59
//   volint           kernel/volint.m:25-263
//   facegeom         kernel/facegeom.m:13-52
//   faceint          kernel/faceint.m:28-360
//   Euler physics    application/FM/euler/{flux,fhat,fbou,getan}.m (2D),
//                    application/FM/euler/{flux,fhat,fbou}_euler3d.c (3D)
//   NS physics       application/FM/ns/{flux2d,fhat,fbou,getan}.m, laplacian AV
//                    application/FM/nsNew/laplacianAV_2d.c (f += eps_av * q)
//   schur_u          kernel/schur_u.c:42-163 (== schur_primal.m:13-49)
//   schur_uq         kernel/schur_uq.c:40-287 (Cholesky of M, dgesv of D)
//   schur_uq2        kernel/schur_primal.m:52-150 (inv(M) form), for cross-checks
//   hdg_elemental    kernel/hdg_elemental.m:1-203
// The dense kernels are its own: unblocked LU with partial pivoting (the
// dgetf2/dgetrs algorithm behind dgesv) and upper Cholesky (dpotrf/dpotrs 'U').
// Physics Jacobians are formed by forward-mode dual numbers; this is the same
// chain rule the reference expands by hand (e.g. euler/getan.m, value plus
// m1..m4 derivative lines) or generates symbolically (ns/flux2d.m).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/hdgx.h"

namespace {

// ---------------------------------------------------------------- dual numbers
template <int N>
struct Dn {
  double v;
  double d[N];
  Dn() : v(0.0) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
  Dn(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
  static Dn var(double x, int i) { Dn r(x); r.d[i] = 1.0; return r; }
};
template <int N> Dn<N> operator+(const Dn<N>& a, const Dn<N>& b) { Dn<N> r; r.v = a.v + b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> Dn<N> operator-(const Dn<N>& a, const Dn<N>& b) { Dn<N> r; r.v = a.v - b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> Dn<N> operator-(const Dn<N>& a) { Dn<N> r; r.v = -a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
template <int N> Dn<N> operator*(const Dn<N>& a, const Dn<N>& b) { Dn<N> r; r.v = a.v * b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> Dn<N> operator/(const Dn<N>& a, const Dn<N>& b) { Dn<N> r; r.v = a.v / b.v; double ib = 1.0 / b.v; for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib; return r; }
template <int N> Dn<N> operator+(const Dn<N>& a, double b) { Dn<N> r = a; r.v += b; return r; }
template <int N> Dn<N> operator+(double b, const Dn<N>& a) { return a + b; }
template <int N> Dn<N> operator-(const Dn<N>& a, double b) { Dn<N> r = a; r.v -= b; return r; }
template <int N> Dn<N> operator-(double b, const Dn<N>& a) { Dn<N> r = -a; r.v += b; return r; }
template <int N> Dn<N> operator*(const Dn<N>& a, double b) { Dn<N> r; r.v = a.v * b; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b; return r; }
template <int N> Dn<N> operator*(double b, const Dn<N>& a) { return a * b; }
template <int N> Dn<N> operator/(const Dn<N>& a, double b) { return a * (1.0 / b); }
template <int N> Dn<N> operator/(double a, const Dn<N>& b) { return Dn<N>(a) / b; }
template <int N> Dn<N> dsqrt(const Dn<N>& a) { Dn<N> r; r.v = std::sqrt(a.v); double h = 0.5 / r.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * h; return r; }
// |x| with MATLAB's sign(0) = 0 for the derivative
template <int N> Dn<N> dabs(const Dn<N>& a) { double s = (a.v > 0) ? 1.0 : ((a.v < 0) ? -1.0 : 0.0); Dn<N> r; r.v = std::fabs(a.v); for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i]; return r; }

constexpr int MAXNC = 20;  // u and q components (NS 2D: 12)
constexpr int MAXU = 5;

// ------------------------------------------------------------------- physics
// Inviscid flux f[c][d] of the Euler equations (euler/flux.m:17-33, flux_euler3d.c)
template <class T>
void euler_flux(int nd, double gam, const T* u, T f[MAXU][3]) {
  const double gam1 = gam - 1.0;
  T r1 = 1.0 / u[0];
  T vel[3];
  T af(0.0);
  for (int i = 0; i < nd; ++i) { vel[i] = u[1 + i] * r1; af = af + 0.5 * vel[i] * vel[i]; }
  T E = u[nd + 1] * r1;
  T p = gam1 * (u[nd + 1] - u[0] * af);
  T h = E + p * r1;
  for (int d = 0; d < nd; ++d) {
    f[0][d] = u[1 + d];
    for (int i = 0; i < nd; ++i) f[1 + i][d] = (i == d) ? u[1 + i] * vel[d] + p : u[1 + i] * vel[d];
    f[nd + 1][d] = u[1 + d] * h;
  }
}

// Navier-Stokes 2D flux with q = (rx,rux,rvx,rEx, ry,ruy,rvy,rEy) in udg[4..11]
// (ns/flux2d.m): f = f_inv + viscous stresses and heat flux written in q.
template <class T>
void ns2d_flux(const double* param, const T* udg, T f[MAXU][3]) {
  const double gam = param[0], gam1 = gam - 1.0, Re = param[2], Pr = param[3];
  const double Re1 = 1.0 / Re, c23 = 2.0 / 3.0, fc = gam / (gam1 * Re * Pr);
  const T& r = udg[0]; const T& ru = udg[1]; const T& rv = udg[2]; const T& rE = udg[3];
  const T& rx = udg[4]; const T& rux = udg[5]; const T& rvx = udg[6]; const T& rEx = udg[7];
  const T& ry = udg[8]; const T& ruy = udg[9]; const T& rvy = udg[10]; const T& rEy = udg[11];
  T r1 = 1.0 / r;
  T uu = ru * r1, vv = rv * r1, E = rE * r1;
  T q2 = 0.5 * (uu * uu + vv * vv);
  T p = gam1 * (rE - r * q2);
  T h = E + p * r1;
  T ux = (rux - rx * uu) * r1, vx = (rvx - rx * vv) * r1;
  T uy = (ruy - ry * uu) * r1, vy = (rvy - ry * vv) * r1;
  T qx = uu * ux + vv * vx, qy = uu * uy + vv * vy;
  T px = gam1 * (rEx - rx * q2 - r * qx), py = gam1 * (rEy - ry * q2 - r * qy);
  T Tx = (px * r - p * rx) * r1 * r1, Ty = (py * r - p * ry) * r1 * r1;
  T txx = Re1 * c23 * (2.0 * ux - vy), txy = Re1 * (uy + vx), tyy = Re1 * c23 * (2.0 * vy - ux);
  f[0][0] = ru;                        f[0][1] = rv;
  f[1][0] = ru * uu + p + txx;         f[1][1] = ru * vv + txy;
  f[2][0] = rv * uu + txy;             f[2][1] = rv * vv + p + tyy;
  f[3][0] = ru * h + uu * txx + vv * txy + fc * Tx;
  f[3][1] = rv * h + uu * txy + vv * tyy + fc * Ty;
}

// Roe-form normal Jacobian An (absflag 0) or |An| with entropy fix (absflag 1):
// ns/getan.m:148-410, FM/getan.c:330-785.  An = rl3*I + a*cc1' + e*cc2'.
template <class T>
void roe_an(int nd, double gam, double eps, int absflag, const T* uh, const double* n, T An[MAXU][MAXU]) {
  const double gam1 = gam - 1.0;
  const int nc = nd + 2;
  T r = uh[0];
  T r1 = 1.0 / r;
  T vel[3];
  T af(0.0), un(0.0);
  for (int i = 0; i < nd; ++i) { vel[i] = uh[1 + i] * r1; af = af + 0.5 * vel[i] * vel[i]; un = un + vel[i] * n[i]; }
  T E = uh[nd + 1] * r1;
  T p = gam1 * (uh[nd + 1] - r * af);
  T h = E + p * r1;
  T c2 = gam * p * r1;
  T c = dsqrt(c2);
  T rl1, rl2, rl3;
  if (absflag) {
    rl1 = dabs(un + c); rl2 = dabs(un - c); rl3 = dabs(un);
    if (eps > 0) {
      T ec = eps * c;
      if (rl1.v < ec.v) rl1 = 0.5 * (rl1 * rl1 / ec + ec);
      if (rl2.v < ec.v) rl2 = 0.5 * (rl2 * rl2 / ec + ec);
      if (rl3.v < ec.v) rl3 = 0.5 * (rl3 * rl3 / ec + ec);
    }
  } else {
    rl1 = un + c; rl2 = un - c; rl3 = un;
  }
  T s1 = 0.5 * (rl1 + rl2), s2 = 0.5 * (rl1 - rl2);
  T a[MAXU], e[MAXU], b[MAXU], dd[MAXU];
  a[0] = T(1.0); e[0] = T(0.0); b[0] = af; dd[0] = -un;
  for (int i = 0; i < nd; ++i) { a[1 + i] = vel[i]; e[1 + i] = T(n[i]); b[1 + i] = -vel[i]; dd[1 + i] = T(n[i]); }
  a[nd + 1] = h; e[nd + 1] = un; b[nd + 1] = T(1.0); dd[nd + 1] = T(0.0);
  T sm = s1 - rl3;
  for (int k = 0; k < nc; ++k) {
    T cc1 = gam1 * sm * b[k] / c2 + s2 * dd[k] / c;
    T cc2 = gam1 * s2 * b[k] / c + sm * dd[k];
    for (int i = 0; i < nc; ++i) {
      T v = cc1 * a[i] + cc2 * e[i];
      An[i][k] = (i == k) ? v + rl3 : v;
    }
  }
}

// Euler 2D local Lax-Friedrichs diagonals (euler/getan.m:25-141): absolute 2
// gives |un|+c, absolute 3 gives un-c.
// With eps > 0 (getan.m:97-103) |un|+c becomes 0.5 (l^2/(eps c) + eps c) and
// its derivative lines use that new value where l is written:
//   l' <- lnew l'/(eps c) + 0.5 (1 - lnew^2/(eps c)^2) eps c'
// (the chain rule would take the old l there); restated as written.
template <int N>
void llf_lam(double gam, double eps, const Dn<N>* uh, const double* n, Dn<N>& lam2, Dn<N>& lam3) {
  typedef Dn<N> T;
  const double gam1 = gam - 1.0;
  T r1 = 1.0 / uh[0];
  T uv = uh[1] * r1, vv = uh[2] * r1;
  T af = 0.5 * (uv * uv + vv * vv);
  T p = gam1 * (uh[3] - uh[0] * af);
  T c = dsqrt(gam * p * r1);
  T un = uv * n[0] + vv * n[1];
  lam2 = dabs(un) + c;
  lam3 = un - c;
  if (eps > 0) {
    const double ec = eps * c.v;
    T l = lam2;
    l.v = 0.5 * (lam2.v * lam2.v / ec + ec);
    for (int i = 0; i < N; ++i)
      l.d[i] = l.v * lam2.d[i] / ec + 0.5 * (1.0 - l.v * l.v / (ec * ec)) * (eps * c.d[i]);
    lam2 = l;
  }
}

struct Phys {
  int physics, nd, ncu, nc, ncd, av_col;
  const double* param;
  int nparam;
  double tau() const { return param[nparam - 1]; }
};

// f (ncu, nd) and f_udg (ncu, nd, nc) at one point
void flux_point(const Phys& P, const double* pg, const double* udg, double* f, double* f_udg) {
  const int ncu = P.ncu, nd = P.nd, nc = P.nc;
  if (P.physics == HDGX_PHYS_EULER) {
    Dn<MAXU> u[MAXU];
    for (int c = 0; c < ncu; ++c) u[c] = Dn<MAXU>::var(udg[c], c);
    Dn<MAXU> fl[MAXU][3];
    euler_flux(nd, P.param[0], u, fl);
    for (int c = 0; c < ncu; ++c)
      for (int d = 0; d < nd; ++d) {
        f[c + ncu * d] = fl[c][d].v;
        for (int k = 0; k < nc; ++k) f_udg[c + ncu * d + ncu * nd * k] = (k < ncu) ? fl[c][d].d[k] : 0.0;
      }
  } else {
    Dn<12> u[12];
    for (int c = 0; c < 12; ++c) u[c] = Dn<12>::var(udg[c], c);
    Dn<12> fl[MAXU][3];
    ns2d_flux(P.param, u, fl);
    double av = (P.av_col >= 0) ? pg[P.av_col] : 0.0;
    for (int c = 0; c < ncu; ++c)
      for (int d = 0; d < nd; ++d) {
        f[c + ncu * d] = fl[c][d].v + av * udg[ncu + c + ncu * d];
        for (int k = 0; k < nc; ++k)
          f_udg[c + ncu * d + ncu * nd * k] = fl[c][d].d[k] + ((k == ncu + c + ncu * d) ? av : 0.0);
      }
  }
}

// Numerical trace flux (euler/fhat.m:21-58 with caseFlux 1, caseStab 4;
// ns/fhat.m:49-67): fh = f(uh[,q]).n + tau (u - uh).
void fhat_point(const Phys& P, const double* nl, const double* pg, const double* udg, const double* uh,
                double* fh, double* fh_udg, double* fh_uh) {
  const int ncu = P.ncu, nd = P.nd, nc = P.nc;
  const double tau = P.tau();
  for (int i = 0; i < ncu * nc; ++i) fh_udg[i] = 0.0;
  if (P.physics == HDGX_PHYS_EULER) {
    Dn<MAXU> u[MAXU];
    for (int c = 0; c < ncu; ++c) u[c] = Dn<MAXU>::var(uh[c], c);
    Dn<MAXU> fl[MAXU][3];
    euler_flux(nd, P.param[0], u, fl);
    for (int c = 0; c < ncu; ++c) {
      Dn<MAXU> fn(0.0);
      for (int d = 0; d < nd; ++d) fn = fn + fl[c][d] * nl[d];
      fh[c] = fn.v + tau * (udg[c] - uh[c]);
      fh_udg[c + ncu * c] = tau;
      for (int k = 0; k < ncu; ++k) fh_uh[c + ncu * k] = fn.d[k] - ((c == k) ? tau : 0.0);
    }
  } else {
    Dn<12> w[12];
    for (int c = 0; c < ncu; ++c) w[c] = Dn<12>::var(uh[c], c);
    for (int c = ncu; c < nc; ++c) w[c] = Dn<12>::var(udg[c], c);
    Dn<12> fl[MAXU][3];
    ns2d_flux(P.param, w, fl);
    double av = (P.av_col >= 0) ? pg[P.av_col] : 0.0;
    for (int c = 0; c < ncu; ++c) {
      Dn<12> fn(0.0);
      for (int d = 0; d < nd; ++d) fn = fn + (fl[c][d] + av * w[ncu + c + ncu * d]) * nl[d];
      fh[c] = fn.v + tau * (udg[c] - uh[c]);
      fh_udg[c + ncu * c] = tau;
      for (int k = ncu; k < nc; ++k) fh_udg[c + ncu * k] = fn.d[k];
      for (int k = 0; k < ncu; ++k) fh_uh[c + ncu * k] = fn.d[k] - ((c == k) ? tau : 0.0);
    }
  }
}

// Boundary flux (euler/fbou.m:21-92, ns/fbou.m:37-101, fbou_euler3d.c:20-60).
// Returns false for a boundary model this oracle does not restate.
bool fbou_point(const Phys& P, int ib, const double* uinf, const double* nl, const double* pg,
                const double* udg, const double* uh, double* fh, double* fh_udg, double* fh_uh) {
  const int ncu = P.ncu, nd = P.nd, nc = P.nc;
  for (int i = 0; i < ncu * nc; ++i) fh_udg[i] = 0.0;
  const bool euler2 = P.physics == HDGX_PHYS_EULER && nd == 2, ns = P.physics == HDGX_PHYS_NS;
  // the far-field form (an + An)(u - uh) - (an - An)(ub - uh) with ub = uinf
  // (ib 1) or the prescribed-pressure state of u: ub = (r, ru, rv, pb/(gam-1)
  // + |ru|^2/(2 r)), pb = uinf(1) (euler ib 3) or (uinf(4) - 0.5)(gam - 1)
  // (euler ib 4, euler/fbou.m:50-77; ns ib 4, ns/fbou.m:76-90)
  const bool pres = nd == 2 && ((euler2 && (ib == 3 || ib == 4)) || (ns && ib == 4));
  if (ib == 1 || pres) {
    double ub[MAXU], Ub[MAXU][MAXU];           // ub and dub/du
    for (int i = 0; i < ncu; ++i) {
      ub[i] = pres ? udg[i] : uinf[i];
      for (int k = 0; k < ncu; ++k) Ub[i][k] = (pres && i == k) ? 1.0 : 0.0;
    }
    if (pres) {
      const double gam = P.param[0], r = udg[0], ru = udg[1], rv = udg[2];
      const double pb = ib == 3 ? uinf[0] : (uinf[3] - 0.5) * (gam - 1.0);
      ub[3] = pb / (gam - 1.0) + 0.5 * (ru * ru + rv * rv) / r;
      Ub[3][0] = -0.5 * (ru * ru + rv * rv) / (r * r);
      Ub[3][1] = ru / r;
      Ub[3][2] = rv / r;
      Ub[3][3] = 0.0;
    }
    Dn<MAXU> w[MAXU];
    for (int c = 0; c < ncu; ++c) w[c] = Dn<MAXU>::var(uh[c], c);
    Dn<MAXU> an[MAXU][MAXU], An[MAXU][MAXU];
    if (euler2) {
      Dn<MAXU> l2, l3;
      llf_lam(P.param[0], P.param[1], w, nl, l2, l3);
      for (int i = 0; i < ncu; ++i)
        for (int k = 0; k < ncu; ++k) { an[i][k] = Dn<MAXU>((i == k) ? 0.0 : 0.0); An[i][k] = an[i][k]; }
      for (int i = 0; i < ncu; ++i) { an[i][i] = l2; An[i][i] = l3; }
    } else {
      roe_an(nd, P.param[0], P.param[1], 0, w, nl, an);
      roe_an(nd, P.param[0], P.param[1], 1, w, nl, An);
    }
    for (int i = 0; i < ncu; ++i) {
      Dn<MAXU> s(0.0);
      for (int k = 0; k < ncu; ++k)
        s = s + (an[i][k] + An[i][k]) * (udg[k] - w[k]) - (an[i][k] - An[i][k]) * (ub[k] - w[k]);
      fh[i] = s.v;
      for (int k = 0; k < ncu; ++k) {
        double v = an[i][k].v + An[i][k].v;
        for (int j = 0; j < ncu; ++j) v -= (an[i][j].v - An[i][j].v) * Ub[j][k];
        fh_udg[i + ncu * k] = v;
        fh_uh[i + ncu * k] = s.d[k];
      }
    }
    return true;
  }
  if (ib == 2 && P.physics == HDGX_PHYS_EULER && nd == 2) {  // inviscid wall, euler/fbou.m:30-48
    double un = udg[1] * nl[0] + udg[2] * nl[1];
    double ub[4] = {udg[0], udg[1] - nl[0] * un, udg[2] - nl[1] * un, udg[3]};
    for (int c = 0; c < 4; ++c) fh[c] = ub[c] - uh[c];
    for (int i = 0; i < 16; ++i) fh_uh[i] = 0.0;
    fh_udg[0 + 4 * 0] = 1.0;
    fh_udg[1 + 4 * 1] = 1.0 - nl[0] * nl[0];
    fh_udg[1 + 4 * 2] = -nl[0] * nl[1];
    fh_udg[2 + 4 * 1] = -nl[0] * nl[1];
    fh_udg[2 + 4 * 2] = 1.0 - nl[1] * nl[1];
    fh_udg[3 + 4 * 3] = 1.0;
    for (int c = 0; c < 4; ++c) fh_uh[c + 4 * c] = -1.0;
    return true;
  }
  for (int i = 0; i < ncu * ncu; ++i) fh_uh[i] = 0.0;
  if (ib == 5 && euler2) {  // euler/fbou.m:78-91: ub = (r, uinf(2:3), rE) of u
    fh[0] = udg[0] - uh[0];
    fh[1] = uinf[1] - uh[1];
    fh[2] = uinf[2] - uh[2];
    fh[3] = udg[3] - uh[3];
    fh_udg[0 + 4 * 0] = 1.0;
    fh_udg[3 + 4 * 3] = 1.0;
    for (int c = 0; c < 4; ++c) fh_uh[c + 4 * c] = -1.0;
    return true;
  }
  if (ib == 2 && ns) {  // adiabatic wall, ns/fbou.m:47-71: r - rh, -ruh, -rvh, fhat's energy row
    double fa[MAXU], fa_udg[MAXU * MAXNC], fa_uh[MAXU * MAXU];
    fhat_point(P, nl, pg, udg, uh, fa, fa_udg, fa_uh);
    fh[0] = udg[0] - uh[0];
    fh[1] = -uh[1];
    fh[2] = -uh[2];
    fh[3] = fa[3];
    fh_udg[0] = 1.0;
    for (int k = 0; k < nc; ++k) fh_udg[3 + ncu * k] = fa_udg[3 + ncu * k];
    for (int c = 0; c < 3; ++c) fh_uh[c + ncu * c] = -1.0;
    for (int k = 0; k < ncu; ++k) fh_uh[3 + ncu * k] = fa_uh[3 + ncu * k];
    return true;
  }
  if (ib == 3 && ns) {  // isothermal wall, ns/fbou.m:72-95: energy row gam (gam-1) M^2 rEh/rh - uinf(end)
    const double gam = P.param[0], M2 = P.param[4] * P.param[4], k = gam * (gam - 1.0) * M2;
    fh[0] = udg[0] - uh[0];
    fh[1] = -uh[1];
    fh[2] = -uh[2];
    fh[3] = k * uh[3] / uh[0] - uinf[ncu - 1];
    fh_udg[0] = 1.0;
    for (int c = 0; c < 3; ++c) fh_uh[c + ncu * c] = -1.0;
    fh_uh[3 + ncu * 0] = -k * uh[3] / (uh[0] * uh[0]);
    fh_uh[3 + ncu * 3] = k / uh[0];
    return true;
  }
  if (ib == 6 && ns) {  // Dirichlet, ns/fbou.m:105-109: uh - uinf
    for (int c = 0; c < ncu; ++c) {
      fh[c] = uh[c] - uinf[c];
      fh_uh[c + ncu * c] = 1.0;
    }
    return true;
  }
  return false;
}

// ---------------------------------------------------------------- dense LA
// LU with partial pivoting of A (n x n, column-major, lda n) and solve for
// B (n x nrhs): the dgetf2 + dgetrs algorithm.  Returns false on a zero pivot.
bool lu_solve(int n, double* A, int nrhs, double* B) {
  std::vector<int> ipiv(n);
  bool ok = true;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double amax = std::fabs(A[k + n * k]);
    for (int i = k + 1; i < n; ++i) {
      double v = std::fabs(A[i + n * k]);
      if (v > amax) { amax = v; p = i; }
    }
    ipiv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(A[k + n * j], A[p + n * j]);
    double piv = A[k + n * k];
    if (piv == 0.0) { ok = false; continue; }
    double ip = 1.0 / piv;
    for (int i = k + 1; i < n; ++i) A[i + n * k] *= ip;
    for (int j = k + 1; j < n; ++j) {
      double u = A[k + n * j];
      if (u != 0.0)
        for (int i = k + 1; i < n; ++i) A[i + n * j] -= A[i + n * k] * u;
    }
  }
  for (int j = 0; j < nrhs; ++j) {
    double* b = B + (size_t)n * j;
    for (int k = 0; k < n; ++k) if (ipiv[k] != k) std::swap(b[k], b[ipiv[k]]);
    for (int k = 0; k < n; ++k) { double v = b[k]; if (v != 0.0) for (int i = k + 1; i < n; ++i) b[i] -= A[i + n * k] * v; }
    for (int k = n - 1; k >= 0; --k) { b[k] /= A[k + n * k]; double v = b[k]; if (v != 0.0) for (int i = 0; i < k; ++i) b[i] -= A[i + n * k] * v; }
  }
  return ok;
}

// Upper Cholesky M = U'U in place (dpotrf 'U'); false if not SPD.
bool chol_upper(int n, double* A) {
  for (int j = 0; j < n; ++j) {
    double s = A[j + n * j];
    for (int k = 0; k < j; ++k) s -= A[k + n * j] * A[k + n * j];
    if (!(s > 0.0)) return false;
    double ujj = std::sqrt(s);
    A[j + n * j] = ujj;
    for (int i = j + 1; i < n; ++i) {
      double t = A[j + n * i];
      for (int k = 0; k < j; ++k) t -= A[k + n * j] * A[k + n * i];
      A[j + n * i] = t / ujj;
    }
  }
  return true;
}

// Solve U'U X = B (dpotrs 'U'), B (n x nrhs) in place.
void chol_solve(int n, const double* U, int nrhs, double* B) {
  for (int j = 0; j < nrhs; ++j) {
    double* b = B + (size_t)n * j;
    for (int i = 0; i < n; ++i) { double s = b[i]; for (int k = 0; k < i; ++k) s -= U[k + n * i] * b[k]; b[i] = s / U[i + n * i]; }
    for (int i = n - 1; i >= 0; --i) { double s = b[i]; for (int k = i + 1; k < n; ++k) s -= U[i + n * k] * b[k]; b[i] = s / U[i + n * i]; }
  }
}

// C (m x n) = alpha*A (m x k) * B (k x n) + beta*C, all column-major
void gemm(int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb,
          double beta, double* C, int ldc) {
  for (int j = 0; j < n; ++j) {
    double* c = C + (size_t)ldc * j;
    if (beta == 0.0) for (int i = 0; i < m; ++i) c[i] = 0.0;
    else if (beta != 1.0) for (int i = 0; i < m; ++i) c[i] *= beta;
    for (int l = 0; l < k; ++l) {
      double b = alpha * B[l + (size_t)ldb * j];
      if (b == 0.0) continue;
      const double* a = A + (size_t)lda * l;
      for (int i = 0; i < m; ++i) c[i] += a[i] * b;
    }
  }
}

bool all_finite(const double* x, size_t n) {
  for (size_t i = 0; i < n; ++i) if (!std::isfinite(x[i])) return false;
  return true;
}

// ------------------------------------------------------------ schur kernels
struct SchurDims { int nd, npv, ncu, nc, ndf; };

// schur_u.c:42-80 localsolve + :143-161 AE/FE, on copies of the inputs.
int schur_u_elem(const SchurDims& s, const double* BD, const double* F, const double* GK, const double* H,
                 const double* Ru, const double* Rh, double* dudg, double* dudg_duh, double* ae, double* fe,
                 std::vector<double>& wk) {
  const int n = s.npv * s.ncu, nh = s.ncu * s.ndf, ncol = s.npv * s.nc;
  wk.resize((size_t)n * n + (size_t)n * (1 + nh));
  double* D = wk.data();
  double* Rt = D + (size_t)n * n;
  for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) D[i + (size_t)n * j] = BD[i + (size_t)n * j];
  for (int i = 0; i < n; ++i) Rt[i] = Ru[i];
  for (int j = 0; j < nh; ++j) for (int i = 0; i < n; ++i) Rt[(size_t)(j + 1) * n + i] = -F[(size_t)j * n + i];
  int st = 0;
  if (!lu_solve(n, D, 1 + nh, Rt)) st |= HDGX_ST_SINGULAR_D;
  // dudg (npv, nc): du in the first npv*ncu entries (nc == ncu for schur_u)
  for (int i = 0; i < s.npv * s.nc; ++i) dudg[i] = (i < n) ? Rt[i] : 0.0;
  for (int j = 0; j < nh; ++j)
    for (int i = 0; i < s.npv * s.nc; ++i) dudg_duh[i + (size_t)ncol * j] = (i < n) ? Rt[(size_t)(j + 1) * n + i] : 0.0;
  for (int i = 0; i < nh * nh; ++i) ae[i] = H[i];
  for (int i = 0; i < nh; ++i) fe[i] = Rh[i];
  gemm(nh, nh, ncol, 1.0, GK, nh, dudg_duh, ncol, 1.0, ae, nh);
  gemm(nh, 1, ncol, -1.0, GK, nh, dudg, ncol, 1.0, fe, nh);
  if (!all_finite(ae, (size_t)nh * nh) || !all_finite(dudg_duh, (size_t)ncol * nh)) st |= HDGX_ST_NONFINITE;
  return st;
}

// schur_uq.c:40-191 localsolveuq + :263-286 AE/FE, on copies of the inputs.
int schur_uq_elem(const SchurDims& s, const double* M0, const double* C0, const double* E0, const double* BD0,
                  const double* F0, const double* GK, const double* H, const double* Rq0, const double* Ru0,
                  const double* Rh, double* dudg, double* dudg_duh, double* ae, double* fe, std::vector<double>& wk) {
  const int nd = s.nd, npv = s.npv, ncu = s.ncu, nc = s.nc, ndf = s.ndf, nd1 = nd + 1;
  const int nu = npv * ncu, nh = ncu * ndf, ncol = npv * nc;
  size_t need = (size_t)npv * npv + (size_t)npv * npv * nd + (size_t)npv * ndf * nd + (size_t)nu * ncol +
                (size_t)nu * nh + (size_t)npv * ncu * nd + nu + nu + (size_t)nu * nu + (size_t)nu * nh +
                (size_t)nu * (1 + nh) + (size_t)npv * ncu * nd + (size_t)npv * ncu * ndf * ncu * nd;
  wk.resize(need);
  double* M = wk.data();
  double* C = M + (size_t)npv * npv;
  double* E = C + (size_t)npv * npv * nd;
  double* BD = E + (size_t)npv * ndf * nd;
  double* F = BD + (size_t)nu * ncol;
  double* Rq = F + (size_t)nu * nh;
  double* Ru = Rq + (size_t)npv * ncu * nd;
  double* BMiR = Ru + nu;
  double* BMiC = BMiR + nu;
  double* BMiE = BMiC + (size_t)nu * nu;
  double* Rt = BMiE + (size_t)nu * nh;
  double* MiCdu = Rt + (size_t)nu * (1 + nh);
  double* MiCdlu = MiCdu + (size_t)npv * ncu * nd;
  std::memcpy(M, M0, sizeof(double) * npv * npv);
  std::memcpy(C, C0, sizeof(double) * npv * npv * nd);
  std::memcpy(E, E0, sizeof(double) * npv * ndf * nd);
  std::memcpy(BD, BD0, sizeof(double) * nu * ncol);
  std::memcpy(F, F0, sizeof(double) * nu * nh);
  std::memcpy(Rq, Rq0, sizeof(double) * npv * ncu * nd);
  std::memcpy(Ru, Ru0, sizeof(double) * nu);
  int st = 0;
  double* D = BD;
  double* B = BD + (size_t)nu * nu;
  if (!chol_upper(npv, M)) st |= HDGX_ST_NONSPD_M;
  chol_solve(npv, M, ncu * nd, Rq);    // Rq <- M^-1 Rq
  chol_solve(npv, M, npv * nd, C);     // C  <- M^-1 C
  chol_solve(npv, M, ndf * nd, E);     // E  <- M^-1 E
  // BMiR = B * vec(M^-1 Rq)   (B is nu x (npv*ncu*nd))
  gemm(nu, 1, npv * ncu * nd, 1.0, B, nu, Rq, npv * ncu * nd, 0.0, BMiR, nu);
  for (int j = 0; j < ncu; ++j)
    for (int i = 0; i < nd; ++i) {
      const double* Bij = B + (size_t)i * nu * nu + (size_t)j * nu * npv;
      double beta = (i == 0) ? 0.0 : 1.0;
      gemm(nu, npv, npv, 1.0, Bij, nu, C + (size_t)i * npv * npv, npv, beta, BMiC + (size_t)j * nu * npv, nu);
      gemm(nu, ndf, npv, 1.0, Bij, nu, E + (size_t)i * npv * ndf, npv, beta, BMiE + (size_t)j * nu * ndf, nu);
    }
  for (int i = 0; i < nu * nu; ++i) D[i] += BMiC[i];
  for (int i = 0; i < nu; ++i) Ru[i] -= BMiR[i];
  // F <- BMiE (re-laid out) - F   (schur_uq.c:115-122)
  for (int j = 0; j < ndf; ++j)
    for (int k = 0; k < ncu; ++k)
      for (int m = 0; m < ncu; ++m)
        for (int n = 0; n < npv; ++n) {
          size_t fi = n + (size_t)npv * m + (size_t)nu * k + (size_t)nu * ncu * j;
          F[fi] = BMiE[n + (size_t)npv * m + (size_t)nu * j + (size_t)nu * ndf * k] - F[fi];
        }
  for (int i = 0; i < nu; ++i) Rt[i] = Ru[i];
  for (int j = 0; j < nh; ++j) for (int i = 0; i < nu; ++i) Rt[(size_t)(j + 1) * nu + i] = F[(size_t)j * nu + i];
  // D is nu x nu with leading dimension nu (first nu columns of BD)
  if (!lu_solve(nu, D, 1 + nh, Rt)) st |= HDGX_ST_SINGULAR_D;
  double* du = Rt;
  double* dlu = Rt + nu;
  for (int i = 0; i < nd; ++i) {
    gemm(npv, ncu, npv, 1.0, C + (size_t)i * npv * npv, npv, du, npv, 0.0, MiCdu + (size_t)i * npv * ncu, npv);
    gemm(npv, ncu * ndf * ncu, npv, 1.0, C + (size_t)i * npv * npv, npv, dlu, npv, 0.0,
         MiCdlu + (size_t)i * npv * ncu * ncu * ndf, npv);
  }
  for (int i = 0; i < nu; ++i) dudg[i] = du[i];
  for (int i = 0; i < npv * ncu * nd; ++i) dudg[nu + i] = Rq[i] + MiCdu[i];
  // dudg_duh (npv, nc, ncu, ndf): u rows from dlu, q rows from MiC*dlu - MiE (x) I
  const size_t s2 = (size_t)npv * ncu * nd1, s3 = s2 * ncu;
  for (int j = 0; j < ndf; ++j)
    for (int k = 0; k < ncu; ++k)
      for (int m = 0; m < ncu; ++m)
        for (int n = 0; n < npv; ++n)
          dudg_duh[n + (size_t)npv * m + s2 * k + s3 * j] = dlu[n + (size_t)npv * m + (size_t)nu * k + (size_t)nu * ncu * j];
  for (int j = 0; j < ndf; ++j)
    for (int k = 0; k < ncu; ++k)
      for (int i = 0; i < nd; ++i)
        for (int m = 0; m < ncu; ++m)
          for (int n = 0; n < npv; ++n) {
            double fac = (m == k) ? 1.0 : 0.0;
            dudg_duh[n + (size_t)npv * m + (size_t)nu * (i + 1) + s2 * k + s3 * j] =
                MiCdlu[n + (size_t)npv * m + (size_t)nu * k + (size_t)nu * ncu * j + (size_t)nu * ncu * ndf * i] -
                fac * E[n + (size_t)npv * j + (size_t)npv * ndf * i];
          }
  for (int i = 0; i < nh * nh; ++i) ae[i] = H[i];
  for (int i = 0; i < nh; ++i) fe[i] = Rh[i];
  gemm(nh, nh, ncol, 1.0, GK, nh, dudg_duh, ncol, 1.0, ae, nh);
  gemm(nh, 1, ncol, -1.0, GK, nh, dudg, ncol, 1.0, fe, nh);
  if (!all_finite(ae, (size_t)nh * nh) || !all_finite(dudg_duh, (size_t)ncol * nh)) st |= HDGX_ST_NONFINITE;
  return st;
}

// schur_primal.m:52-150 (schur_uq2): the same elimination through inv(M).
int schur_uq2_elem(const SchurDims& s, const double* M0, const double* C0, const double* E0, const double* BD0,
                   const double* F0, const double* GK, const double* H, const double* Rq0, const double* Ru0,
                   const double* Rh, double* dudg, double* dudg_duh, double* ae, double* fe) {
  const int nd = s.nd, npv = s.npv, ncu = s.ncu, nc = s.nc, ndf = s.ndf;
  const int nu = npv * ncu, nh = ncu * ndf, ncol = npv * nc;
  std::vector<double> Mi((size_t)npv * npv, 0.0), Mc(M0, M0 + npv * npv);
  for (int i = 0; i < npv; ++i) Mi[i + (size_t)npv * i] = 1.0;
  int st = 0;
  if (!lu_solve(npv, Mc.data(), npv, Mi.data())) st |= HDGX_ST_NONSPD_M;   // Mi = inv(M)
  std::vector<double> MiRq((size_t)npv * ncu * nd), MiC((size_t)npv * npv * nd), MiE((size_t)npv * ndf * nd);
  gemm(npv, ncu * nd, npv, 1.0, Mi.data(), npv, Rq0, npv, 0.0, MiRq.data(), npv);
  for (int d = 0; d < nd; ++d) {
    gemm(npv, npv, npv, 1.0, Mi.data(), npv, C0 + (size_t)d * npv * npv, npv, 0.0, MiC.data() + (size_t)d * npv * npv, npv);
    gemm(npv, ndf, npv, 1.0, Mi.data(), npv, E0 + (size_t)d * npv * ndf, npv, 0.0, MiE.data() + (size_t)d * npv * ndf, npv);
  }
  // B(:,:,:,j,d) blocks; BMiR = B * MiRq(:); BMiC(:,:,:,j) = sum_d B(:,:,:,j,d) MiC(:,:,d)
  const double* B = BD0 + (size_t)nu * nu;
  std::vector<double> BMiR(nu), SD(BD0, BD0 + (size_t)nu * nu), BMiE((size_t)nu * ncu * ndf, 0.0);
  gemm(nu, 1, npv * ncu * nd, 1.0, B, nu, MiRq.data(), npv * ncu * nd, 0.0, BMiR.data(), nu);
  for (int j = 0; j < ncu; ++j)
    for (int d = 0; d < nd; ++d) {
      const double* Bjd = B + (size_t)(j + ncu * d) * nu * npv;
      gemm(nu, npv, npv, 1.0, Bjd, nu, MiC.data() + (size_t)d * npv * npv, npv, 1.0, SD.data() + (size_t)j * nu * npv, nu);
      // BMiE (npv,ncu, ncu(j), ndf) column (j + ncu*b)
      std::vector<double> tmp((size_t)nu * ndf);
      gemm(nu, ndf, npv, 1.0, Bjd, nu, MiE.data() + (size_t)d * npv * ndf, npv, 0.0, tmp.data(), nu);
      for (int b = 0; b < ndf; ++b)
        for (int i = 0; i < nu; ++i) BMiE[i + (size_t)nu * (j + ncu * b)] += tmp[i + (size_t)nu * b];
    }
  std::vector<double> rhs((size_t)nu * (1 + nh));
  for (int i = 0; i < nu; ++i) rhs[i] = Ru0[i] - BMiR[i];
  for (int j = 0; j < nh; ++j) for (int i = 0; i < nu; ++i) rhs[(size_t)nu * (1 + j) + i] = BMiE[i + (size_t)nu * j] - F0[i + (size_t)nu * j];
  if (!lu_solve(nu, SD.data(), 1 + nh, rhs.data())) st |= HDGX_ST_SINGULAR_D;
  const double* du = rhs.data();
  const double* dlu = rhs.data() + nu;
  for (int i = 0; i < nu; ++i) dudg[i] = du[i];
  for (int d = 0; d < nd; ++d) {
    std::vector<double> t((size_t)npv * ncu);
    gemm(npv, ncu, npv, 1.0, MiC.data() + (size_t)d * npv * npv, npv, du, npv, 0.0, t.data(), npv);
    for (int i = 0; i < npv * ncu; ++i) dudg[nu + (size_t)d * nu + i] = MiRq[(size_t)d * nu + i] + t[i];
  }
  for (int b = 0; b < nh; ++b) {
    const double* x = dlu + (size_t)nu * b;
    double* out = dudg_duh + (size_t)ncol * b;
    for (int i = 0; i < nu; ++i) out[i] = x[i];
    int k = b % ncu, jf = b / ncu;
    for (int d = 0; d < nd; ++d) {
      std::vector<double> t((size_t)nu);
      gemm(npv, ncu, npv, 1.0, MiC.data() + (size_t)d * npv * npv, npv, x, npv, 0.0, t.data(), npv);
      for (int m = 0; m < ncu; ++m)
        for (int n = 0; n < npv; ++n)
          out[nu + (size_t)d * nu + n + (size_t)npv * m] =
              t[n + (size_t)npv * m] - ((m == k) ? MiE[n + (size_t)npv * jf + (size_t)npv * ndf * d] : 0.0);
    }
  }
  for (int i = 0; i < nh * nh; ++i) ae[i] = H[i];
  for (int i = 0; i < nh; ++i) fe[i] = Rh[i];
  gemm(nh, nh, ncol, 1.0, GK, nh, dudg_duh, ncol, 1.0, ae, nh);
  gemm(nh, 1, ncol, -1.0, GK, nh, dudg, ncol, 1.0, fe, nh);
  return st;
}

// ------------------------------------------------- volint + faceint per element
struct ElemOut {
  double *BD, *F, *GK, *H, *Ru, *Rh, *M, *C, *E, *Rq;
};

bool elem_matrices(const hdgx_master* ms, const hdgx_app* ap, const Phys& P, const double* X,
                   const int32_t* bfe, const double* U, const double* UH, const double* SH, ElemOut o,
                   std::vector<double>& wk) {
  const int nd = ms->nd, npv = ms->npv, npf = ms->npf, nfe = ms->nfe, ngv = ms->ngv, ngf = ms->ngf;
  const int ncu = ap->ncu, nc = ap->nc, ncd = ap->ncd, nd1 = nd + 1, ndf = npf * nfe;
  const int nu = npv * ncu;
  const double *shapvt = ms->shapvt, *shapvg = ms->shapvg, *svv = ms->shapvgdotshapvl;
  const double *shapft = ms->shapft, *shapfg = ms->shapfg, *sff = ms->shapfgdotshapfc;
  const double *shapmv = ms->shapmv, *shapmf = ms->shapmf;
  const int32_t *perm = ms->perm, *permg = ms->permgeom;
  const bool ls = ap->localsolve != 0;

  size_t need = (size_t)ngv * (ncd + 1 + nd * nd + nc + ncu * nd + ncu * nd * nc + ncu + ncu * nc) +
                (size_t)ngv * nd1 * (ncu + ncu * nc) + (size_t)ngf * (ncd + nd + 1 + nc + ncu) +
                (size_t)ngf * 2 * (ncu + ncu * nc + ncu * ncu) + 64;
  wk.resize(need);
  double* p = wk.data();
  double* pg = p; p += (size_t)ngv * ncd;
  double* jac = p; p += ngv;
  double* Xx = p; p += (size_t)ngv * nd * nd;           // Xx[g][a][b]
  double* ug = p; p += (size_t)ngv * nc;
  double* f = p; p += (size_t)ngv * ncu * nd;          // f[g][(c,d)]
  double* fu = p; p += (size_t)ngv * ncu * nd * nc;    // f_udg[g][(c,d,k)]
  double* s = p; p += (size_t)ngv * ncu;
  double* su = p; p += (size_t)ngv * ncu * nc;
  double* wrk = p; p += (size_t)ngv * nd1 * ncu;       // wrk[(g + ngv*i)][c]  column-major (ngv*nd1, ncu)
  double* wrl = p; p += (size_t)ngv * nd1 * ncu * nc;  // (ngv*nd1, ncu*nc)
  double* fpg = p; p += (size_t)ngf * ncd;
  double* nl = p; p += (size_t)ngf * nd;
  double* fjac = p; p += ngf;
  double* fug = p; p += (size_t)ngf * nc;
  double* fuh = p; p += (size_t)ngf * ncu;
  double* FH = p; p += (size_t)ngf * ncu;
  double* FHu = p; p += (size_t)ngf * ncu * nc;
  double* FHh = p; p += (size_t)ngf * ncu * ncu;
  double* BH = p; p += (size_t)ngf * ncu;
  double* BHu = p; p += (size_t)ngf * ncu * nc;
  double* BHh = p; p += (size_t)ngf * ncu * ncu;

  auto Xv = [&](int a, int q) { return X[a + (size_t)npv * q]; };
  // ---- volgeom
  for (int g = 0; g < ngv; ++g) {
    for (int q = 0; q < ncd; ++q) {
      double v = 0;
      for (int a = 0; a < npv; ++a) v += shapmv[g + (size_t)ngv * a] * Xv(a, q);
      pg[g + (size_t)ngv * q] = v;
    }
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int j = 0; j < nd; ++j)
      for (int i = 0; i < nd; ++i) {
        double v = 0;
        for (int a = 0; a < npv; ++a) v += shapmv[g + (size_t)ngv * a + (size_t)ngv * npv * (1 + j)] * Xv(a, i);
        J[j][i] = v;   // Jg(:,:,j,i) = dx_i/dxi_j
      }
    double* xx = Xx + (size_t)g * nd * nd;  // xx[a*nd+b] = Xx(:,:,a,b)
    if (nd == 2) {
      jac[g] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      xx[0 * 2 + 0] = J[1][1];
      xx[1 * 2 + 0] = -J[1][0];
      xx[0 * 2 + 1] = -J[0][1];
      xx[1 * 2 + 1] = J[0][0];
    } else {
      jac[g] = J[0][0] * J[1][1] * J[2][2] - J[0][0] * J[2][1] * J[1][2] + J[1][0] * J[2][1] * J[0][2] -
               J[1][0] * J[0][1] * J[2][2] + J[2][0] * J[0][1] * J[1][2] - J[2][0] * J[1][1] * J[0][2];
      xx[0 * 3 + 0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      xx[1 * 3 + 0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      xx[2 * 3 + 0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      xx[0 * 3 + 1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
      xx[1 * 3 + 1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
      xx[2 * 3 + 1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
      xx[0 * 3 + 2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
      xx[1 * 3 + 2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
      xx[2 * 3 + 2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    }
  }
  // ---- volint
  for (int g = 0; g < ngv; ++g)
    for (int c = 0; c < nc; ++c) {
      double v = 0;
      for (int a = 0; a < npv; ++a) v += shapvt[g + (size_t)ngv * a] * U[a + (size_t)npv * c];
      ug[g + (size_t)ngv * c] = v;
    }
  if (ls) {
    for (int ii = 0; ii < npv; ++ii)
      for (int jj = 0; jj < npv; ++jj) {
        double v = 0;
        for (int g = 0; g < ngv; ++g) v += svv[jj + (size_t)npv * ii + (size_t)npv * npv * g] * jac[g];
        o.M[jj + (size_t)npv * ii] = ap->fc_q * v;
      }
    for (int i = 0; i < nd; ++i)
      for (int ii = 0; ii < npv; ++ii)
        for (int jj = 0; jj < npv; ++jj) {
          double v = 0;
          for (int j = 0; j < nd; ++j)
            for (int g = 0; g < ngv; ++g)
              v += svv[jj + (size_t)npv * ii + (size_t)npv * npv * g + (size_t)npv * npv * ngv * (j + 1)] *
                   Xx[(size_t)g * nd * nd + i * nd + j];
          o.C[jj + (size_t)npv * ii + (size_t)npv * npv * i] = v;
        }
    for (int i = 0; i < nd; ++i)
      for (int c = 0; c < ncu; ++c)
        for (int a = 0; a < npv; ++a) {
          double v = 0;
          const int qc = ncu + c + ncu * i;
          for (int b = 0; b < npv; ++b)
            v += o.M[a + (size_t)npv * b] * (SH[b + (size_t)npv * qc] / ap->fc_q - U[b + (size_t)npv * qc]);
          double w = 0;
          for (int b = 0; b < npv; ++b) w += o.C[a + (size_t)npv * b + (size_t)npv * npv * i] * U[b + (size_t)npv * c];
          o.Rq[a + (size_t)npv * c + (size_t)npv * ncu * i] = v + w;
        }
  }
  for (int g = 0; g < ngv; ++g) {
    double up[MAXNC], pp[4], fp[MAXU * 3], fup[MAXU * 3 * MAXNC];
    for (int c = 0; c < nc; ++c) up[c] = ug[g + (size_t)ngv * c];
    for (int q = 0; q < ncd; ++q) pp[q] = pg[g + (size_t)ngv * q];
    flux_point(P, pp, up, fp, fup);
    for (int i = 0; i < ncu * nd; ++i) f[(size_t)g * ncu * nd + i] = fp[i];
    for (int i = 0; i < ncu * nd * nc; ++i) fu[(size_t)g * ncu * nd * nc + i] = fup[i];
    for (int i = 0; i < ncu; ++i) s[(size_t)g * ncu + i] = 0.0;          // zero source (euler/source.m, ns/source.m)
    for (int i = 0; i < ncu * nc; ++i) su[(size_t)g * ncu * nc + i] = 0.0;
    if (ap->tdep) {
      for (int c = 0; c < ncu; ++c) {
        double st = 0;
        for (int a = 0; a < npv; ++a) st += shapvt[g + (size_t)ngv * a] * SH[a + (size_t)npv * c];
        s[(size_t)g * ncu + c] += st - up[c] * ap->fc_u;
        su[(size_t)g * ncu * nc + c + (size_t)ncu * c] -= ap->fc_u;
      }
    }
  }
  const int nr = ngv * nd1;
  for (int g = 0; g < ngv; ++g) {
    for (int c = 0; c < ncu; ++c) {
      wrk[g + (size_t)nr * c] = s[(size_t)g * ncu + c] * jac[g];
      for (int k = 0; k < nc; ++k)
        wrl[g + (size_t)nr * (c + ncu * k)] = -su[(size_t)g * ncu * nc + c + (size_t)ncu * k] * jac[g];
    }
    const double* xx = Xx + (size_t)g * nd * nd;
    for (int i = 0; i < nd; ++i) {
      for (int c = 0; c < ncu; ++c) {
        double v = 0;
        for (int j = 0; j < nd; ++j) v += f[(size_t)g * ncu * nd + c + ncu * j] * xx[j * nd + i];
        wrk[g + (size_t)ngv * (i + 1) + (size_t)nr * c] = v;
        for (int k = 0; k < nc; ++k) {
          double w = 0;
          for (int j = 0; j < nd; ++j) w += fu[(size_t)g * ncu * nd * nc + c + ncu * j + (size_t)ncu * nd * k] * xx[j * nd + i];
          wrl[g + (size_t)ngv * (i + 1) + (size_t)nr * (c + ncu * k)] = -w;
        }
      }
    }
  }
  // Ru = shapvg * wrk ; BD = shapvgdotshapvl * wrl  (volint.m:258-262)
  for (int c = 0; c < ncu; ++c)
    for (int a = 0; a < npv; ++a) {
      double v = 0;
      for (int r = 0; r < nr; ++r) v += shapvg[a + (size_t)npv * r] * wrk[r + (size_t)nr * c];
      o.Ru[a + (size_t)npv * c] = v;
    }
  for (int k = 0; k < nc; ++k)
    for (int c = 0; c < ncu; ++c)
      for (int ii = 0; ii < npv; ++ii)
        for (int jj = 0; jj < npv; ++jj) {
          double v = 0;
          for (int r = 0; r < nr; ++r) v += svv[jj + (size_t)npv * ii + (size_t)npv * npv * r] * wrl[r + (size_t)nr * (c + ncu * k)];
          o.BD[jj + (size_t)npv * c + (size_t)nu * ii + (size_t)nu * npv * k] = v;
        }
  // ---- faceint
  const int nh = ncu * ndf;
  for (int i = 0; i < nu * nh; ++i) o.F[i] = 0.0;
  for (int i = 0; i < nh * npv * nc; ++i) o.GK[i] = 0.0;
  for (int i = 0; i < nh * nh; ++i) o.H[i] = 0.0;
  if (ls) for (int i = 0; i < npv * ndf * nd; ++i) o.E[i] = 0.0;
  bool ok = true;
  for (int j = 0; j < nfe; ++j) {
    // facegeom
    for (int g = 0; g < ngf; ++g) {
      for (int q = 0; q < ncd; ++q) {
        double v = 0;
        for (int a = 0; a < npf; ++a) v += shapmf[g + (size_t)ngf * a] * Xv(permg[a + npf * j], q);
        fpg[g + (size_t)ngf * q] = v;
      }
      double dpg[3][2];
      for (int d = 0; d < nd; ++d)
        for (int k = 0; k < nd - 1; ++k) {
          double v = 0;
          for (int a = 0; a < npf; ++a)
            v += shapmf[g + (size_t)ngf * a + (size_t)ngf * npf * (1 + k)] * Xv(permg[a + npf * j], d);
          dpg[d][k] = v;
        }
      double n0, n1, n2 = 0, jj;
      if (nd == 2) {
        jj = std::sqrt(dpg[0][0] * dpg[0][0] + dpg[1][0] * dpg[1][0]);
        n0 = dpg[1][0] / jj;
        n1 = -dpg[0][0] / jj;
      } else {
        n0 = dpg[1][0] * dpg[2][1] - dpg[2][0] * dpg[1][1];
        n1 = dpg[2][0] * dpg[0][1] - dpg[0][0] * dpg[2][1];
        n2 = dpg[0][0] * dpg[1][1] - dpg[1][0] * dpg[0][1];
        jj = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        n0 /= jj; n1 /= jj; n2 /= jj;
      }
      fjac[g] = jj;
      nl[g * nd + 0] = n0;
      nl[g * nd + 1] = n1;
      if (nd == 3) nl[g * nd + 2] = n2;
    }
    if (ls) {  // E (faceint.m:34-47) and Rq -= E*UH (:48-51)
      for (int i = 0; i < nd; ++i)
        for (int b = 0; b < npf; ++b)
          for (int a = 0; a < npf; ++a) {
            double v = 0;
            for (int g = 0; g < ngf; ++g) v += sff[a + (size_t)npf * b + (size_t)npf * npf * g] * nl[g * nd + i] * fjac[g];
            o.E[perm[a + npf * j] + (size_t)npv * (b + npf * j) + (size_t)npv * ndf * i] = v;
          }
    }
    for (int g = 0; g < ngf; ++g) {
      for (int c = 0; c < nc; ++c) {
        double v = 0;
        for (int a = 0; a < npf; ++a) v += shapft[g + (size_t)ngf * a] * U[perm[a + npf * j] + (size_t)npv * c];
        fug[g + (size_t)ngf * c] = v;
      }
      for (int c = 0; c < ncu; ++c) {
        double v = 0;
        for (int a = 0; a < npf; ++a) v += shapft[g + (size_t)ngf * a] * UH[c + (size_t)ncu * (a + npf * j)];
        fuh[g + (size_t)ngf * c] = v;
      }
    }
    const bool bnd = bfe[j] < 0;
    int ib = 0;
    const double* uinf = nullptr;
    if (bnd) {
      int id = -bfe[j];
      if (id < 1 || id > ap->nbc) return false;
      ib = ap->bcm[id - 1];
      uinf = ap->bcs + (size_t)(id - 1) * HDGX_MAX_PARAM;
    }
    for (int g = 0; g < ngf; ++g) {
      double up[MAXNC], uhp[MAXU], pp[4], nn[3];
      for (int c = 0; c < nc; ++c) up[c] = fug[g + (size_t)ngf * c];
      for (int c = 0; c < ncu; ++c) uhp[c] = fuh[g + (size_t)ngf * c];
      for (int q = 0; q < ncd; ++q) pp[q] = fpg[g + (size_t)ngf * q];
      for (int d = 0; d < nd; ++d) nn[d] = nl[g * nd + d];
      fhat_point(P, nn, pp, up, uhp, FH + (size_t)g * ncu, FHu + (size_t)g * ncu * nc, FHh + (size_t)g * ncu * ncu);
      if (bnd && !fbou_point(P, ib, uinf, nn, pp, up, uhp, BH + (size_t)g * ncu, BHu + (size_t)g * ncu * nc,
                             BHh + (size_t)g * ncu * ncu))
        ok = false;
    }
    // Run, BDn, Fn (fhat) into Ru, BD, F; Rh, Gt, Ht (fbou on boundary faces) into Rh, GK, H
    for (int c = 0; c < ncu; ++c)
      for (int a = 0; a < npf; ++a) {
        double v = 0, w = 0;
        for (int g = 0; g < ngf; ++g) {
          double sj = shapfg[a + (size_t)npf * g] * fjac[g];
          v += sj * FH[(size_t)g * ncu + c];
          if (bnd) w += sj * BH[(size_t)g * ncu + c];
        }
        o.Ru[perm[a + npf * j] + (size_t)npv * c] -= v;
        o.Rh[c + (size_t)ncu * (a + npf * j)] = -(bnd ? w : v);
      }
    for (int b = 0; b < npf; ++b)
      for (int a = 0; a < npf; ++a) {
        const int va = perm[a + npf * j], vb = perm[b + npf * j];
        for (int c = 0; c < ncu; ++c) {
          for (int k = 0; k < nc; ++k) {
            double v = 0, w = 0;
            for (int g = 0; g < ngf; ++g) {
              double sj = sff[a + (size_t)npf * b + (size_t)npf * npf * g] * fjac[g];
              v += sj * FHu[(size_t)g * ncu * nc + c + (size_t)ncu * k];
              if (bnd) w += sj * BHu[(size_t)g * ncu * nc + c + (size_t)ncu * k];
            }
            o.BD[va + (size_t)npv * c + (size_t)nu * vb + (size_t)nu * npv * k] += v;
            o.GK[c + (size_t)ncu * (b + npf * j) + (size_t)nh * va + (size_t)nh * npv * k] += bnd ? w : v;
          }
          for (int k = 0; k < ncu; ++k) {
            double v = 0, w = 0;
            for (int g = 0; g < ngf; ++g) {
              double sj = sff[a + (size_t)npf * b + (size_t)npf * npf * g] * fjac[g];
              v += sj * FHh[(size_t)g * ncu * ncu + c + (size_t)ncu * k];
              if (bnd) w += sj * BHh[(size_t)g * ncu * ncu + c + (size_t)ncu * k];
            }
            o.F[va + (size_t)npv * c + (size_t)nu * k + (size_t)nu * ncu * (b + npf * j)] += v;
            o.H[c + (size_t)ncu * (a + npf * j) + (size_t)nh * k + (size_t)nh * ncu * (b + npf * j)] += bnd ? w : v;
          }
        }
      }
  }
  if (ls) {
    for (int i = 0; i < nd; ++i)
      for (int c = 0; c < ncu; ++c)
        for (int a = 0; a < npv; ++a) {
          double v = 0;
          for (int b = 0; b < ndf; ++b)
            v += o.E[a + (size_t)npv * b + (size_t)npv * ndf * i] * UH[c + (size_t)ncu * b];
          o.Rq[a + (size_t)npv * c + (size_t)npv * ncu * i] -= v;
        }
  }
  return ok;
}

Phys make_phys(const hdgx_app* ap) {
  Phys P;
  P.physics = ap->physics; P.nd = ap->nd; P.ncu = ap->ncu; P.nc = ap->nc; P.ncd = ap->ncd;
  P.av_col = ap->av_col; P.param = ap->param; P.nparam = ap->nparam;
  return P;
}

bool check_app(const hdgx_master* ms, const hdgx_app* ap) {
  if (!ms || !ap) return false;
  if (ap->nd != ms->nd) return false;
  if (ap->physics == HDGX_PHYS_EULER && ap->ncu != ms->nd + 2) return false;
  if (ap->physics == HDGX_PHYS_NS && (ms->nd != 2 || ap->ncu != 4 || ap->nc != 12)) return false;
  if (ap->localsolve && ap->nc != ap->ncu * (ms->nd + 1)) return false;
  if (!ap->localsolve && ap->nc != ap->ncu) return false;
  return true;
}

int nthreads_or_default(int nt) {
#ifdef _OPENMP
  return nt > 0 ? nt : omp_get_max_threads();
#else
  (void)nt;
  return 1;
#endif
}

}  // namespace

extern "C" {

// Per-point physics, point-fastest layouts as the reference's (ng, ...) arrays.
int oracle_flux(const hdgx_app* ap, int ng, const double* pg, const double* udg, double* f, double* f_udg) {
  Phys P = make_phys(ap);
  const int ncu = P.ncu, nd = P.nd, nc = P.nc, ncd = P.ncd;
  for (int g = 0; g < ng; ++g) {
    double pp[4], up[MAXNC], fp[MAXU * 3], fup[MAXU * 3 * MAXNC];
    for (int q = 0; q < ncd; ++q) pp[q] = pg[g + (size_t)ng * q];
    for (int c = 0; c < nc; ++c) up[c] = udg[g + (size_t)ng * c];
    flux_point(P, pp, up, fp, fup);
    for (int i = 0; i < ncu * nd; ++i) f[g + (size_t)ng * i] = fp[i];
    for (int i = 0; i < ncu * nd * nc; ++i) f_udg[g + (size_t)ng * i] = fup[i];
  }
  return 0;
}

int oracle_fhat(const hdgx_app* ap, int ng, const double* nl, const double* pg, const double* udg,
                const double* uh, double* fh, double* fh_udg, double* fh_uh) {
  Phys P = make_phys(ap);
  const int ncu = P.ncu, nd = P.nd, nc = P.nc, ncd = P.ncd;
  for (int g = 0; g < ng; ++g) {
    double pp[4], up[MAXNC], uhp[MAXU], nn[3], a[MAXU], b[MAXU * MAXNC], c[MAXU * MAXU];
    for (int q = 0; q < ncd; ++q) pp[q] = pg[g + (size_t)ng * q];
    for (int i = 0; i < nc; ++i) up[i] = udg[g + (size_t)ng * i];
    for (int i = 0; i < ncu; ++i) uhp[i] = uh[g + (size_t)ng * i];
    for (int d = 0; d < nd; ++d) nn[d] = nl[g + (size_t)ng * d];
    fhat_point(P, nn, pp, up, uhp, a, b, c);
    for (int i = 0; i < ncu; ++i) fh[g + (size_t)ng * i] = a[i];
    for (int i = 0; i < ncu * nc; ++i) fh_udg[g + (size_t)ng * i] = b[i];
    for (int i = 0; i < ncu * ncu; ++i) fh_uh[g + (size_t)ng * i] = c[i];
  }
  return 0;
}

int oracle_fbou(const hdgx_app* ap, int ib, const double* uinf, int ng, const double* nl, const double* pg,
                const double* udg, const double* uh, double* fh, double* fh_udg, double* fh_uh) {
  Phys P = make_phys(ap);
  const int ncu = P.ncu, nd = P.nd, nc = P.nc, ncd = P.ncd;
  for (int g = 0; g < ng; ++g) {
    double pp[4], up[MAXNC], uhp[MAXU], nn[3], a[MAXU], b[MAXU * MAXNC], c[MAXU * MAXU];
    for (int q = 0; q < ncd; ++q) pp[q] = pg[g + (size_t)ng * q];
    for (int i = 0; i < nc; ++i) up[i] = udg[g + (size_t)ng * i];
    for (int i = 0; i < ncu; ++i) uhp[i] = uh[g + (size_t)ng * i];
    for (int d = 0; d < nd; ++d) nn[d] = nl[g + (size_t)ng * d];
    if (!fbou_point(P, ib, uinf, nn, pp, up, uhp, a, b, c)) return HDGX_ERR_UNSUPPORTED;
    for (int i = 0; i < ncu; ++i) fh[g + (size_t)ng * i] = a[i];
    for (int i = 0; i < ncu * nc; ++i) fh_udg[g + (size_t)ng * i] = b[i];
    for (int i = 0; i < ncu * ncu; ++i) fh_uh[g + (size_t)ng * i] = c[i];
  }
  return 0;
}

int oracle_elemental_matrices(const hdgx_master* ms, const hdgx_app* ap, int64_t ne, const double* dgnodes,
                              const int32_t* bf, const double* udg, const double* uh, const double* sh, double* BD,
                              double* F, double* GK, double* H, double* Ru, double* Rh, double* M, double* C,
                              double* E, double* Rq, int nthreads) {
  if (!check_app(ms, ap)) return HDGX_ERR_ARG;
  const int npv = ms->npv, ncu = ap->ncu, nc = ap->nc, nd = ms->nd, ncd = ap->ncd, nfe = ms->nfe;
  const int ndf = ms->npf * nfe, nu = npv * ncu, nh = ncu * ndf;
  const bool ls = ap->localsolve != 0;
  Phys P = make_phys(ap);
  int err = 0;
#pragma omp parallel num_threads(nthreads_or_default(nthreads))
  {
    std::vector<double> wk, zeros((size_t)npv * nc, 0.0);
#pragma omp for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
      ElemOut o;
      o.BD = BD + (size_t)e * nu * npv * nc;
      o.F = F + (size_t)e * nu * nh;
      o.GK = GK + (size_t)e * nh * npv * nc;
      o.H = H + (size_t)e * nh * nh;
      o.Ru = Ru + (size_t)e * nu;
      o.Rh = Rh + (size_t)e * nh;
      o.M = ls ? M + (size_t)e * npv * npv : nullptr;
      o.C = ls ? C + (size_t)e * npv * npv * nd : nullptr;
      o.E = ls ? E + (size_t)e * npv * ndf * nd : nullptr;
      o.Rq = ls ? Rq + (size_t)e * npv * ncu * nd : nullptr;
      const double* she = sh ? sh + (size_t)e * npv * nc : zeros.data();
      if (!elem_matrices(ms, ap, P, dgnodes + (size_t)e * npv * ncd, bf + (size_t)e * nfe,
                         udg + (size_t)e * npv * nc, uh + (size_t)e * nh, she, o, wk)) {
#pragma omp atomic write
        err = HDGX_ERR_UNSUPPORTED;
      }
    }
  }
  return err;
}

int oracle_schur_u(int npv, int ncu, int nc, int ndf, int64_t ne, const double* BD, const double* F,
                   const double* GK, const double* H, const double* Ru, const double* Rh, double* dudg,
                   double* dudg_duh, double* ae, double* fe, int32_t* status, int nthreads) {
  if (nc != ncu) return HDGX_ERR_ARG;
  SchurDims s{0, npv, ncu, nc, ndf};
  const size_t nu = (size_t)npv * ncu, nh = (size_t)ncu * ndf, ncol = (size_t)npv * nc;
#pragma omp parallel num_threads(nthreads_or_default(nthreads))
  {
    std::vector<double> wk;
#pragma omp for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
      int st = schur_u_elem(s, BD + e * nu * ncol, F + e * nu * nh, GK + e * nh * ncol, H + e * nh * nh,
                            Ru + e * nu, Rh + e * nh, dudg + e * ncol, dudg_duh + e * ncol * nh, ae + e * nh * nh,
                            fe + e * nh, wk);
      if (status) status[e] = st;
    }
  }
  return 0;
}

int oracle_schur_uq(int nd, int npv, int ncu, int nc, int ndf, int64_t ne, const double* M, const double* C,
                    const double* E, const double* BD, const double* F, const double* GK, const double* H,
                    const double* Rq, const double* Ru, const double* Rh, double* dudg, double* dudg_duh,
                    double* ae, double* fe, int32_t* status, int nthreads) {
  if (nc != ncu * (nd + 1)) return HDGX_ERR_ARG;
  SchurDims s{nd, npv, ncu, nc, ndf};
  const size_t nu = (size_t)npv * ncu, nh = (size_t)ncu * ndf, ncol = (size_t)npv * nc;
#pragma omp parallel num_threads(nthreads_or_default(nthreads))
  {
    std::vector<double> wk;
#pragma omp for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
      int st = schur_uq_elem(s, M + e * npv * npv, C + e * npv * npv * nd, E + e * npv * ndf * nd,
                             BD + e * nu * ncol, F + e * nu * nh, GK + e * nh * ncol, H + e * nh * nh,
                             Rq + e * npv * ncu * nd, Ru + e * nu, Rh + e * nh, dudg + e * ncol,
                             dudg_duh + e * ncol * nh, ae + e * nh * nh, fe + e * nh, wk);
      if (status) status[e] = st;
    }
  }
  return 0;
}

int oracle_schur_uq2(int nd, int npv, int ncu, int nc, int ndf, int64_t ne, const double* M, const double* C,
                     const double* E, const double* BD, const double* F, const double* GK, const double* H,
                     const double* Rq, const double* Ru, const double* Rh, double* dudg, double* dudg_duh,
                     double* ae, double* fe, int32_t* status) {
  if (nc != ncu * (nd + 1)) return HDGX_ERR_ARG;
  SchurDims s{nd, npv, ncu, nc, ndf};
  const size_t nu = (size_t)npv * ncu, nh = (size_t)ncu * ndf, ncol = (size_t)npv * nc;
  for (int64_t e = 0; e < ne; ++e) {
    int st = schur_uq2_elem(s, M + e * npv * npv, C + e * npv * npv * nd, E + e * npv * ndf * nd,
                            BD + e * nu * ncol, F + e * nu * nh, GK + e * nh * ncol, H + e * nh * nh,
                            Rq + e * npv * ncu * nd, Ru + e * nu, Rh + e * nh, dudg + e * ncol,
                            dudg_duh + e * ncol * nh, ae + e * nh * nh, fe + e * nh);
    if (status) status[e] = st;
  }
  return 0;
}

// hdg_elemental.m:1-203: volint + faceint + schur_primal per element, with
// per-thread scratch (the reference processes blocks of 20000 elements; the
// arithmetic per element is the same).
int oracle_elemental(const hdgx_master* ms, const hdgx_app* ap, int64_t ne, const double* dgnodes,
                     const int32_t* bf, const double* udg, const double* uh, const double* sh, double* ae,
                     double* fe, double* dudg, double* dudg_duh, int32_t* status, int nthreads) {
  if (!check_app(ms, ap)) return HDGX_ERR_ARG;
  const int npv = ms->npv, ncu = ap->ncu, nc = ap->nc, nd = ms->nd, ncd = ap->ncd, nfe = ms->nfe;
  const int ndf = ms->npf * nfe, nu = npv * ncu, nh = ncu * ndf, ncol = npv * nc;
  const bool ls = ap->localsolve != 0;
  Phys P = make_phys(ap);
  SchurDims s{nd, npv, ncu, nc, ndf};
  int err = 0;
#pragma omp parallel num_threads(nthreads_or_default(nthreads))
  {
    std::vector<double> wk, wk2, zeros((size_t)npv * nc, 0.0);
    std::vector<double> buf((size_t)nu * ncol + (size_t)nu * nh + (size_t)nh * ncol + (size_t)nh * nh + nu + nh +
                            (size_t)npv * npv + (size_t)npv * npv * nd + (size_t)npv * ndf * nd + (size_t)npv * ncu * nd);
#pragma omp for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
      ElemOut o;
      double* p = buf.data();
      o.BD = p; p += (size_t)nu * ncol;
      o.F = p; p += (size_t)nu * nh;
      o.GK = p; p += (size_t)nh * ncol;
      o.H = p; p += (size_t)nh * nh;
      o.Ru = p; p += nu;
      o.Rh = p; p += nh;
      o.M = p; p += (size_t)npv * npv;
      o.C = p; p += (size_t)npv * npv * nd;
      o.E = p; p += (size_t)npv * ndf * nd;
      o.Rq = p;
      const double* she = sh ? sh + (size_t)e * npv * nc : zeros.data();
      if (!elem_matrices(ms, ap, P, dgnodes + (size_t)e * npv * ncd, bf + (size_t)e * nfe,
                         udg + (size_t)e * npv * nc, uh + (size_t)e * nh, she, o, wk)) {
#pragma omp atomic write
        err = HDGX_ERR_UNSUPPORTED;
      }
      int st;
      if (ls)
        st = schur_uq_elem(s, o.M, o.C, o.E, o.BD, o.F, o.GK, o.H, o.Rq, o.Ru, o.Rh, dudg + (size_t)e * ncol,
                           dudg_duh + (size_t)e * ncol * nh, ae + (size_t)e * nh * nh, fe + (size_t)e * nh, wk2);
      else
        st = schur_u_elem(s, o.BD, o.F, o.GK, o.H, o.Ru, o.Rh, dudg + (size_t)e * ncol,
                          dudg_duh + (size_t)e * ncol * nh, ae + (size_t)e * nh * nh, fe + (size_t)e * nh, wk2);
      if (status) status[e] = st;
    }
  }
  return err;
}

}  // extern "C"
