// Element matrices of the HDG elemental path on gfx950: one workgroup per
// element computes everything volint + faceint produce (kernel/volint.m:1-263,
// kernel/faceint.m:1-360) straight into the layouts the Schur step consumes.
//
// Dataflow per element (each phase a flat task list over the workgroup):
//   0. stage dgnodes, UDG, the gathered UH, SH and the small master tables in
//      LDS;
//   1. pointwise inputs: volume geometry (volgeom.m:22-59), face geometry
//      (facegeom.m:13-52), and the state, trace and SH values at every
//      Gauss point, one task per (point, component);
//   2. physics, one task per (Gauss point, derivative direction): the flux
//      (or fhat / fbou) is evaluated once in single-direction dual numbers,
//      giving one column of its Jacobian; the results are the weighted rows
//      wrk/wrl of volint.m:241-254 and the face rows of faceint.m:55-277
//      pre-multiplied by the surface Jacobian;
//   3. outputs, each task owning a run of consecutive entries: BD one column
//      (all npv test functions), F/GK/H/E/M/C/Rq/Ru/Rh single entries.  The
//      face scatter of faceint.m:322-360 (perm) becomes the lookup loc(v, j)
//      of volume node v on face j, so nothing is read-modify-written.
//
// The volume part of BD uses shapvgdotshapvl in its factored form
// shapvg(jj,g,d) * shapvl(ii,g) (util/preprocess.m:23-37): a BD column
// (c, ii, k) is sum_r shapvg(:, r) * (shapvl(ii, g_r) wrl(r, c, k)), whose
// shapvg operand is uniform across the wavefront.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "hdgx_device.h"
#include "hdgx_kernels.h"
#include "hdgx_shapes.h"

namespace hdgx {

constexpr int QUAD_THREADS = 256;

// Threads of one element's workgroup: 64 where one wavefront stages every
// input array in one pass and takes each task phase (pointwise inputs,
// physics) in at most two rounds, else 256.  The smallest orders then keep
// many elements in flight per CU instead of one 256-thread workgroup that
// idles (2D Euler p = 1: 0.57 -> 0.36 ms for 80 000 elements); at p = 2 and 3
// fewer threads a workgroup measured no faster (one wavefront) or slower (two).
template <class C>
struct QuadThreads {
  static constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }
  static constexpr int mx(int a, int b) { return a > b ? a : b; }
  static constexpr int need() {
    constexpr int NFG = C::NFE * C::NGF;
    int n = mx(C::NPV * (C::ND + 2), C::NPV * C::NC);
    n = mx(n, mx(C::NH, C::NPV * C::NFE));
    n = mx(n, mx(C::NGV * C::NPV, C::NPF * C::NPF * C::NGF));
    n = mx(n, C::NGF * C::NPF * C::ND);
    const int p1 = C::NGV + NFG + C::NGV * C::NC + NFG * C::NC + NFG * C::NCU + C::NGV * C::NCU;
    const int p2 = C::NGV * C::NC + NFG * (C::NC + C::NCU);
    return mx(n, mx(cdiv(p1, 2), cdiv(p2, 2)));
  }
  static constexpr int N = need() <= 64 ? 64 : 256;
};

// wavefronts per SIMD the quadrature kernels are compiled for: the
// single-call kernel four in 2D (NS p = 3 with its gather keys loaded in the
// loops; 40 KB of LDS then allows four workgroups per CU), three for
// tetrahedra; the batched kernel four for 2D Euler (it would otherwise take 8
// accumulator registers more and drop to three), three otherwise
template <class C>
struct QuadWpe {
  static constexpr int N1 = C::ND == 2 ? 4 : 3;
  static constexpr int N = (!C::UQ && C::ND == 2) ? 4 : 3;
};

typedef Dual<1> D1;

// Loads through the constant address space: with a wavefront-uniform index
// they become scalar loads whose result feeds VALU instructions as an SGPR
// operand (for master tables, which no kernel writes).
typedef const __attribute__((address_space(4))) double* ConstD;
typedef const __attribute__((address_space(4))) int32_t* ConstI;
HDGX_DEV double uload(const double* p, int i) { return ((ConstD)p)[i]; }
HDGX_DEV int32_t uload(const int32_t* p, int i) { return ((ConstI)p)[i]; }

// f (NCU x ND) at one point (euler/flux.m, ns/flux2d.m; laplacian AV f += av*q
// as in nsshock/flux.m:16-21)
// Jacobian determinant and adjugate of J[a][b] = dx_b/dxi_a: Xx = det(J) J^-1,
// Xx[a][b] = jac * dxi_b/dx_a (kernel/volgeom.m:22-46).
template <int ND>
HDGX_DEV double adjugate(const double J[ND][ND], double Xx[ND][ND]) {
  if constexpr (ND == 2) {
    Xx[0][0] = J[1][1]; Xx[1][0] = -J[1][0]; Xx[0][1] = -J[0][1]; Xx[1][1] = J[0][0];
    return J[0][0] * J[1][1] - J[0][1] * J[1][0];
  } else {
    Xx[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    Xx[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    Xx[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    Xx[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    Xx[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    Xx[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
    Xx[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    Xx[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    Xx[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    return J[0][0] * J[1][1] * J[2][2] - J[0][0] * J[2][1] * J[1][2] + J[1][0] * J[2][1] * J[0][2] -
           J[1][0] * J[0][1] * J[2][2] + J[2][0] * J[0][1] * J[1][2] - J[2][0] * J[1][1] * J[0][2];
  }
}

// Unit normal and surface Jacobian from the face tangents dpg[d][k] =
// dx_d/ds_k (kernel/facegeom.m:13-52).
template <int ND>
HDGX_DEV double face_normal(const double dpg[ND][ND > 1 ? ND - 1 : 1], double nl[ND]) {
  if constexpr (ND == 2) {
    const double jac = sqrt(dpg[0][0] * dpg[0][0] + dpg[1][0] * dpg[1][0]);
    nl[0] = dpg[1][0] / jac;
    nl[1] = -dpg[0][0] / jac;
    return jac;
  } else {
    const double n0 = dpg[1][0] * dpg[2][1] - dpg[2][0] * dpg[1][1];
    const double n1 = dpg[2][0] * dpg[0][1] - dpg[0][0] * dpg[2][1];
    const double n2 = dpg[0][0] * dpg[1][1] - dpg[1][0] * dpg[0][1];
    const double jac = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    nl[0] = n0 / jac; nl[1] = n1 / jac; nl[2] = n2 / jac;
    return jac;
  }
}

// Volume Gauss point g: jac and Xx from the geometry basis shapmv and the
// element nodes X (npv, ncd), column-major with leading dimension NPV.
template <class C>
HDGX_DEV double vol_geom(const double* shapmv, const double* X, int g, double Xx[C::ND][C::ND]) {
  constexpr int ND = C::ND, NPV = C::NPV, NGV = C::NGV;
  double J[ND][ND];
#pragma unroll
  for (int j = 0; j < ND; ++j)
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < NPV; ++a) v += shapmv[g + NGV * a + NGV * NPV * (1 + j)] * X[a + NPV * i];
      J[j][i] = v;   // dx_i / dxi_j
    }
  return adjugate<ND>(J, Xx);
}

// Face j, Gauss point g: unit normal and surface Jacobian.
template <class C>
HDGX_DEV double face_geom(const double* shapmf, const int32_t* permgeom, const double* X, int j, int g, double nl[C::ND]) {
  constexpr int ND = C::ND, NPV = C::NPV, NPF = C::NPF, NGF = C::NGF;
  double dpg[ND][ND > 1 ? ND - 1 : 1];
#pragma unroll
  for (int d = 0; d < ND; ++d)
#pragma unroll
    for (int k = 0; k < ND - 1; ++k) {
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < NPF; ++a) v += shapmf[g + NGF * a + NGF * NPF * (1 + k)] * X[permgeom[a + NPF * j] + NPV * d];
      dpg[d][k] = v;
    }
  return face_normal<ND>(dpg, nl);
}

template <class C>
HDGX_DEV double vol_geom(const DevMaster& T, const double* X, int g, double Xx[C::ND][C::ND]) {
  return vol_geom<C>(T.shapmv, X, g, Xx);
}
template <class C>
HDGX_DEV double face_geom(const DevMaster& T, const double* X, int j, int g, double nl[C::ND]) {
  return face_geom<C>(T.shapmf, T.permgeom, X, j, g, nl);
}

template <class C, class T>
HDGX_DEV void vol_flux(const DevApp& ap, double av, const T* u, T f[C::NCU][C::ND]) {
  if constexpr (C::PHYS == PHYS_EULER) {
    euler_flux<C::ND>(ap.gam, u, f);
  } else {
    ns2d_flux(ap.gam, ap.Re, ap.Pr, u, u + 4, f);
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int d = 0; d < 2; ++d) f[c][d] = f[c][d] + av * u[4 + c + 4 * d];
  }
}

// fhat = f(uh[, q]).n + tau (u - uh) (euler/fhat.m caseFlux 1 / caseStab 4,
// ns/fhat.m:49-67; AV adds av q.n).  u and uh may differ in type (one of them
// carries the derivative direction, the other is constant).
template <class C, class TU, class TH>
HDGX_DEV void face_fhat(const DevApp& ap, double av, const double* n, const TU* u, const TH* uh,
                        prom_t<TU, TH> fh[C::NCU]) {
  constexpr int ND = C::ND, NCU = C::NCU;
  // the flux of uh (and, with viscosity, of q from u)
  typedef typename std::conditional<C::PHYS == PHYS_EULER, TH, prom_t<TU, TH>>::type TF;
  TF fl[NCU][ND];
  if constexpr (C::PHYS == PHYS_EULER) euler_flux<ND>(ap.gam, uh, fl);
  else ns2d_flux(ap.gam, ap.Re, ap.Pr, uh, u + 4, fl);
#pragma unroll
  for (int c = 0; c < NCU; ++c) {
    TF s = fl[c][0] * n[0];
#pragma unroll
    for (int d = 1; d < ND; ++d) s = s + fl[c][d] * n[d];
    if constexpr (C::PHYS == PHYS_NS) s = s + av * (u[4 + c] * n[0] + u[8 + c] * n[1]);
    fh[c] = s + ap.tau * (u[c] - uh[c]);
  }
}

// Boundary flux.  The far-field form fh = (an + An)(u - uh) - (an - An)(ub - uh)
// = an (u - ub) + An (u + ub - 2 uh), with the LLF pair an = |un| + c,
// An = un - c in 2D Euler (euler/fbou.m:24-29, getan.m; eps > 0 smooths an as
// getan.m:97-103 writes it) and the Roe pair an = A_n, An = |A_n| for NS and
// 3D Euler (ns/fbou.m:37-46, fbou_euler3d.c:20-60), takes ub = uinf for the far
// field (ib 1) and the prescribed-pressure state of u, ub = (r, ru, rv,
// pb/(gam-1) + |ru|^2/(2r)) with pb = uinf(1) (Euler ib 3) or (uinf(4) - 0.5)
// (gam - 1) (Euler and NS ib 4: euler/fbou.m:50-77, ns/fbou.m:76-90).  The
// others: 2D Euler inviscid wall (ib 2, euler/fbou.m:30-48) and inflow (ib 5,
// :78-91); NS adiabatic wall (ib 2, ns/fbou.m:47-71; its energy row is fhat's,
// passed in fa), isothermal wall (ib 3, :72-95) and Dirichlet (ib 6,
// :105-109).  Host validation rejects every other boundary type.
template <class C, class TU, class TH>
HDGX_DEV void face_fbou(const DevApp& ap, int ib, const double* uinf, const double* n, const TU* u, const TH* uh,
                        const prom_t<TU, TH>* fa, prom_t<TU, TH> fh[C::NCU]) {
  constexpr int ND = C::ND, NCU = C::NCU;
  typedef prom_t<TU, TH> T;
  if constexpr (ND == 2) {
    if (C::PHYS == PHYS_EULER && ib == 2) {
      TU un = u[1] * n[0] + u[2] * n[1];
      fh[0] = u[0] - uh[0];
      fh[1] = (u[1] - n[0] * un) - uh[1];
      fh[2] = (u[2] - n[1] * un) - uh[2];
      fh[3] = u[3] - uh[3];
      return;
    }
    if (C::PHYS == PHYS_EULER && ib == 5) {
      fh[0] = u[0] - uh[0];
      fh[1] = uinf[1] - uh[1];
      fh[2] = uinf[2] - uh[2];
      fh[3] = u[3] - uh[3];
      return;
    }
    if (C::PHYS == PHYS_NS && (ib == 2 || ib == 3)) {
      fh[0] = u[0] - uh[0];
      fh[1] = 0.0 - uh[1];
      fh[2] = 0.0 - uh[2];
      if (ib == 2) fh[3] = fa[3];
      else fh[3] = ap.gam * (ap.gam - 1.0) * ap.Minf * ap.Minf * uh[3] / uh[0] - uinf[3];
      return;
    }
    if (C::PHYS == PHYS_NS && ib == 6) {
#pragma unroll
      for (int k = 0; k < NCU; ++k) fh[k] = uh[k] - uinf[k];
      return;
    }
  }
  // the far-field form with ub = uinf or the prescribed-pressure state of u
  TU ub[NCU];
#pragma unroll
  for (int k = 0; k < NCU; ++k) ub[k] = uinf[k];
  if constexpr (ND == 2) {
    if (ib == 3 || ib == 4) {
      const double pb = ib == 3 ? uinf[0] : (uinf[3] - 0.5) * (ap.gam - 1.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) ub[k] = u[k];
      ub[3] = pb / (ap.gam - 1.0) + 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0];
    }
  }
  if constexpr (C::PHYS == PHYS_EULER && ND == 2) {
    const double gam1 = ap.gam - 1.0;
    TH r1 = 1.0 / uh[0];
    TH uv = uh[1] * r1, vv = uh[2] * r1;
    TH af = 0.5 * (uv * uv + vv * vv);
    TH p = gam1 * (uh[3] - uh[0] * af);
    TH c = dsqrt(ap.gam * p * r1);
    TH un = uv * n[0] + vv * n[1];
    TH l2 = dabs(un) + c, l3 = un - c;
    if (ap.eps > 0) l2 = llf_smooth(l2, c, ap.eps);
#pragma unroll
    for (int k = 0; k < NCU; ++k) fh[k] = l2 * (u[k] - ub[k]) + l3 * ((u[k] - uh[k]) + (ub[k] - uh[k]));
  } else {
    TU x0[NCU];
    T x[NCU], y[NCU];
#pragma unroll
    for (int k = 0; k < NCU; ++k) x0[k] = u[k] - ub[k];
    roe_apply<ND>(ap.gam, ap.eps, 0, uh, n, x0, y);
#pragma unroll
    for (int k = 0; k < NCU; ++k) { fh[k] = y[k]; x[k] = (u[k] - uh[k]) + (ub[k] - uh[k]); }
    roe_apply<ND>(ap.gam, ap.eps, 1, uh, n, x, y);
#pragma unroll
    for (int k = 0; k < NCU; ++k) fh[k] = fh[k] + y[k];
  }
}

// one face point's fhat (fa) and boundary flux (fb: fbou on a boundary face,
// else fhat), differentiated along whichever of u, uh is the dual
template <class C, class TU, class TH>
HDGX_DEV void face_point(const DevApp& ap, double av, const double* nl, const TU* u, const TH* uh, int bf,
                         D1 fa[C::NCU], D1 fb[C::NCU]) {
  constexpr int NCU = C::NCU;
  prom_t<TU, TH> fh[NCU];
  face_fhat<C>(ap, av, nl, u, uh, fh);
#pragma unroll
  for (int c = 0; c < NCU; ++c) fa[c] = fh[c];
  if (bf < 0) {
    const int id = -bf - 1;
    double uinf[NCU];
#pragma unroll
    for (int c = 0; c < NCU; ++c) uinf[c] = ap.bcs[id * 5 + c];
    prom_t<TU, TH> fa_[NCU];
#pragma unroll
    for (int c = 0; c < NCU; ++c) fa_[c] = fh[c];
    face_fbou<C>(ap, ap.bcm[id], uinf, nl, u, uh, fa_, fh);
  }
#pragma unroll
  for (int c = 0; c < NCU; ++c) fb[c] = fh[c];
}

// Gather keys of F, GK and H in the compact workspace layouts (WsLayout in
// hdgx_shapes.h), fixed by the configuration alone and built at compile time:
// every entry is a sum over the face points g of shapfgdotshapfc(la, b, g)
// times a face-flux Jacobian entry X(j NGF + g, (c, k)) (faceint.m:316-359),
// so a key carries the two LDS offsets and the output loops do no index
// arithmetic.  -1: a structural zero.
//   face blocks (F and H in both layouts, GK in FbShape's): so | xo << 8,
//     with so = la + NPF b the pair's offset and xo = j NGF + NFG (c + NCU k)
//     the column's;
//   fragment-tile GK: tp | j << 10 | b << 14 | c << 18, the column (va, k)
//     read at run time from the AE k-step table at tp.
// SG (face products staged by FaceStage, fragment tiles only): the sums are
// already in LDS, so F's and H's keys are the staged indices themselves;
// GK's is unchanged.
template <class C, bool FB, bool SG = false>
struct FghKeys {
  using S = S64Shape<C>;
  using FS = FbShape<C>;
  static constexpr int NPV = C::NPV, NPF = C::NPF, NCU = C::NCU, NDF = C::NDF, NH = C::NH, NU = C::NU;
  static constexpr int NGF = C::NGF, NFG = C::NFE * C::NGF, NB = FS::NB;
  static constexpr int NF = C::NFE * NB * NB;    // F's face blocks in both layouts
  static constexpr int NG = FB ? (int)FS::BLK : S::KSP * S::TRH * 64;
  static constexpr int NHK = C::NFE * NB * NB;   // H's face blocks in both layouts
  int32_t f[NF], gk[NG], h[NHK];
};

template <class C, bool FB, bool SG = false>
constexpr FghKeys<C, FB, SG> make_fgh_keys() {
  using K = FghKeys<C, FB, SG>;
  using S = S64Shape<C>;
  constexpr int NPV = K::NPV, NPF = K::NPF, NCU = K::NCU, NDF = K::NDF, NH = K::NH, NB = K::NB;
  constexpr int NGF = K::NGF, NFG = K::NFG;
  K k{};
  auto pair = [](int so, int xo) { return so | (xo << 8); };
  // F_f(q, r): q = a + NPF m (face node a of face f, component m), r = kk + NCU b
  for (int idx = 0; idx < K::NF; ++idx) {
    const int f = idx / (NB * NB), rem = idx % (NB * NB), q = rem % NB, r = rem / NB;
    const int a = q % NPF, m = q / NPF, kk = r % NCU, b = r / NCU;
    if constexpr (SG) k.f[idx] = (f * NPF * NPF + a + NPF * b) * NCU * NCU + m + NCU * kk;
    else k.f[idx] = pair(a + NPF * b, f * NGF + NFG * (m + NCU * kk));
  }
  if constexpr (FB) {
    for (int idx = 0; idx < K::NG; ++idx) {
      // GK_f(r, q): trace row r = c + NCU b, volume q = a + NPF kk
      const int f = idx / (NB * NB), rem = idx % (NB * NB), r = rem % NB, q = rem / NB;
      k.gk[idx] = pair(q % NPF + NPF * (r / NCU), f * NGF + NFG * (r % NCU + NCU * (q / NPF)));
    }
  } else {
    for (int idx = 0; idx < K::NG; ++idx) {
      const int ln = idx & 63, rest = idx >> 6, rta = rest % S::TRH, s = rest / S::TRH;
      const int row = 16 * rta + (ln & 15);
      if (row >= NH) { k.gk[idx] = -1; continue; }
      const int tp = S::TRH + rta * 4 * S::KSMAX + 4 * s + (ln >> 4);
      const int c = row % NCU, bj = row / NCU, j = bj / NPF, b = bj % NPF;
      k.gk[idx] = tp | (j << 10) | (b << 14) | (c << 18);
    }
  }
  // H_f(r, r') = H(f NB + r, f NB + r'): face node a = r / NCU, component
  // c = r % NCU; b, kk likewise from r'
  for (int idx = 0; idx < K::NHK; ++idx) {
    const int f = idx / (NB * NB), rem = idx % (NB * NB), r = rem % NB, r2 = rem / NB;
    const int a = r / NCU, c = r % NCU, b = r2 / NCU, kk = r2 % NCU;
    if constexpr (SG) k.h[idx] = (f * NPF * NPF + a + NPF * b) * NCU * NCU + c + NCU * kk;
    else k.h[idx] = pair(a + NPF * b, f * NGF + NFG * (c + NCU * kk));
  }
  return k;
}

template <class C, bool FB, bool SG = false>
struct FghKeyTab {
  static constexpr FghKeys<C, FB, SG> tab = make_fgh_keys<C, FB, SG>();
};

// The face products of F, GK and H as FP64 MFMA products staged in LDS: per
// face f, P_f(pair, col) = sum_g shapfgdotshapfc(pair, g) X(f, g, col)
// (faceint.m:316-359), pair = a + NPF b the face-node pair, X the fhat
// Jacobian columns col of the array (F: c + NCU k of fhat_u, GK: c + NCU k of
// fhat_udg, H: c + NCU k of fhat_uh).  Staged at (f NPF^2 + pair) NX + col in
// the BD product's operand buffer once BD is done: GK alone, then F and H.
// For the fragment tiles of the 2D orders, where GK's products fit there.
template <class C>
struct FaceStage {
  static constexpr int PT = C::NPF * C::NPF, RTP = (PT + 15) / 16, KSF = (C::NGF + 3) / 4;
  static constexpr int NXF = C::NCU * C::NCU, NXG = C::NCU * C::NC, NXH = C::NCU * C::NCU;
  static constexpr int WORDS_GK = C::NFE * PT * NXG, WORDS_FH = C::NFE * PT * (NXF + NXH);
  static constexpr int CAP = BdShape<C>::KT * BdShape<C>::NCC;
  static constexpr bool USE = S64Shape<C>::USE && !FbShape<C>::USE && WORDS_GK <= CAP && WORDS_FH <= CAP;
};

// LDS of one element's quadrature (one workgroup).  RO: residual only (no
// Jacobian arrays).
template <class C, bool RO = false>
struct QuadLds {
  static constexpr int ND = C::ND, NPV = C::NPV, NPF = C::NPF, NFE = C::NFE, NDF = C::NDF;
  static constexpr int NGV = C::NGV, NGF = C::NGF, NCU = C::NCU, NC = C::NC, NH = C::NH;
  static constexpr int NR = C::NR, NFG = NFE * NGF;
  static constexpr bool UQ = C::UQ;
  double sU[NPV * NC];
  double sUH[NH];
  double sSH[NPV * NC];
  int32_t sBf[NFE];
  int32_t sLoc[NPV * NFE];
  int32_t sPerm[NPF * NFE];
  int32_t sPermG[NPF * NFE];
  unsigned char sAek[RO ? 4 : (S64Shape<C>::TRH * (1 + 4 * S64Shape<C>::KSMAX) + 3) / 4 * 4];
  double sSff[NPF * NPF * NGF];
  double sJac[NGV];
  double sXx[NGV * ND * ND];
  double sAv[NGV];
  double sUg[NGV * NC];
  double sSg[NGV * NCU];
  double sNl[NFG * ND];
  double sFJ[NFG];
  double sAvf[NFG];
  double sUf[NFG * NC];
  double sUHf[NFG * NCU];
  double sWrk[NR * NCU];
  // the B rows of BD's product (BdShape): volume rows wrl (r < NR), then the
  // face rows fhat_udg (NR + t), each row's NCU*NC columns contiguous.  They
  // are first written by the physics phase, so they share their storage with
  // the tables only the staging and pointwise phases read (the element's
  // coordinates, shapmv, shapmf and the interpolation tables).
  struct Ph1 {
    double sX[NPV * (ND + 1)];
    double sShM[NGV * NPV * (ND + 1)];     // shapmv
    double sShMF[NGF * NPF * ND];          // shapmf
    double sShV[NGV * NPV];
    double sShF[NGF * NPF];
  };
  union {
    Ph1 ph1;
    double sWL[RO ? 1 : BdShape<C>::KT * BdShape<C>::NCC + 16];
  };
  double sFH[NFG * NCU];
  double sFHh[RO ? 1 : NFG * NCU * NCU];
  double sGH[NFG * NCU];
  double sGHu[RO ? 1 : NFG * NCU * NC];
  double sGHh[RO ? 1 : NFG * NCU * NCU];
  double sM[UQ ? NPV * NPV : 1];
  double sC[UQ ? NPV * NPV * ND : 1];
  double sE[UQ ? NPV * NDF * ND : 1];
};

// Element e of the batch A (a workgroup of NT_ threads), LDS in L.
// RO: the residual only (hdg_residual.m:1-57: Ru, Rq and Rh with that
// function's sign, +shapfg (fbou | fhat), hdg_residual.m:286-290), no
// Jacobians and no element matrices.
template <class C, bool RO = false, bool TILE = false, int NT_ = QUAD_THREADS>
HDGX_DEV void quad_body(const QuadArgs& A, const DevMaster& T, const DevApp& ap, const int64_t e,
                        QuadLds<C, RO>& L) {
  constexpr int ND = C::ND, NPV = C::NPV, NPF = C::NPF, NFE = C::NFE, NDF = C::NDF;
  constexpr int NGV = C::NGV, NGF = C::NGF, NCU = C::NCU, NC = C::NC, NU = C::NU, NH = C::NH;
  constexpr int NR = C::NR, NFG = NFE * NGF, NT = NT_;
  static_assert(NT % 64 == 0 && NT <= 256, "whole wavefronts");
  constexpr int NCC = NCU * NC;          // BD columns per volume node pair: (c, k)
  constexpr int NDIRF = NC + NCU;        // face derivative directions: UDG components, then UH
  constexpr bool UQ = C::UQ;

  const int tid = threadIdx.x;
  if (e >= A.ne) return;
  const int ncd = ap.ncd;
  const bool tdep = ap.tdep != 0;

  double* const sX = L.ph1.sX;
  double* const sU = L.sU;
  double* const sUH = L.sUH;
  double* const sSH = L.sSH;
  int32_t* const sBf = L.sBf;
  int32_t* const sLoc = L.sLoc;
  double* const sShV = L.ph1.sShV;
  double* const sShF = L.ph1.sShF;
  double* const sSff = L.sSff;
  double* const sJac = L.sJac;
  double* const sXx = L.sXx;
  double* const sAv = L.sAv;
  double* const sUg = L.sUg;
  double* const sSg = L.sSg;
  double* const sNl = L.sNl;
  double* const sFJ = L.sFJ;
  double* const sAvf = L.sAvf;
  double* const sUf = L.sUf;
  double* const sUHf = L.sUHf;
  double* const sWrk = L.sWrk;
  double* const sWL = L.sWL;
  double* const sFH = L.sFH;
  double* const sFHh = L.sFHh;
  double* const sGH = L.sGH;
  double* const sGHu = L.sGHu;
  double* const sGHh = L.sGHh;
  double* const sM = L.sM;
  double* const sC = L.sC;
  double* const sE = L.sE;

  // ---- 0. stage the element's inputs and the small master tables: every
  // array fits one pass of the workgroup, so all loads are issued before
  // the first LDS store (one memory latency instead of one per array)
  {
    static_assert(NPV * (ND + 2) <= NT && NPV * NC <= NT && NH <= NT && NPV * NFE <= NT && NGV * NPV <= NT &&
                  NGF * NPF <= NT && NPF * NPF * NGF <= NT, "staged arrays fit one workgroup pass");
    const bool sh_on = UQ || tdep;
    const double x = tid < NPV * ncd ? A.dgnodes[e * NPV * ncd + tid] : 0.0;
    const double u = tid < NPV * NC ? A.udg[e * NPV * NC + tid] : 0.0;
    const double uh = tid < NH ? A.uh[e * NH + tid] : 0.0;
    const double sh = (sh_on && A.sh && tid < NPV * NC) ? A.sh[e * NPV * NC + tid] : 0.0;
    const int32_t bf = tid < NFE ? A.bf[e * NFE + tid] : 0;
    const int32_t loc = tid < NPV * NFE ? T.loc[tid] : 0;
    const double shv = tid < NGV * NPV ? T.shapvt[tid] : 0.0;
    const double shf = tid < NGF * NPF ? T.shapft[tid] : 0.0;
    const double sff = tid < NPF * NPF * NGF ? T.sff[tid] : 0.0;
    // the geometry tables, node maps and (fragment-tile GK) the AE k-step table
    constexpr int NMV = NGV * NPV * (ND + 1), NMVI = (NMV + NT - 1) / NT;
    constexpr int NAEK = (int)sizeof(L.sAek) / 4;
    static_assert(NGF * NPF * ND <= NT && NAEK <= NT, "staged arrays fit one workgroup pass");
    double mv[NMVI];
#pragma unroll
    for (int it = 0; it < NMVI; ++it) mv[it] = tid + it * NT < NMV ? T.shapmv[tid + it * NT] : 0.0;
    const double mf = tid < NGF * NPF * ND ? T.shapmf[tid] : 0.0;
    const int32_t pm = tid < NPF * NFE ? T.perm[tid] : 0;
    const int32_t pg = tid < NPF * NFE ? T.permgeom[tid] : 0;
    const bool aek_on = TILE && !RO && !FbShape<C>::USE;
    const int32_t ak = (aek_on && tid < NAEK) ? reinterpret_cast<const int32_t*>(T.aek[UQ ? 1 : 0])[tid] : 0;
#pragma unroll
    for (int it = 0; it < NMVI; ++it)
      if (tid + it * NT < NMV) L.ph1.sShM[tid + it * NT] = mv[it];
    if (tid < NGF * NPF * ND) L.ph1.sShMF[tid] = mf;
    if (tid < NPF * NFE) { L.sPerm[tid] = pm; L.sPermG[tid] = pg; }
    if (aek_on && tid < NAEK) reinterpret_cast<int32_t*>(L.sAek)[tid] = ak;
    if (tid < NPV * ncd) sX[tid] = x;
    if (tid < NPV * NC) sU[tid] = u;
    if (tid < NH) sUH[tid] = uh;
    if (sh_on && tid < NPV * NC) sSH[tid] = sh;
    if (tid < NFE) sBf[tid] = bf;
    if (tid < NPV * NFE) sLoc[tid] = loc;
    if (tid < NGV * NPV) sShV[tid] = shv;
    if (tid < NGF * NPF) sShF[tid] = shf;
    if (tid < NPF * NPF * NGF) sSff[tid] = sff;
  }
  __syncthreads();

  // ---- 1. pointwise inputs: the state, trace and SH at every Gauss point
  // (one task per point and component), then the volume and face geometry on
  // the last threads, which the interpolation's last round leaves idle
  {
    constexpr int W1 = NGV * NC, W2 = W1 + NFG * NC, W3 = W2 + NFG * NCU;
    const int W4 = W3 + (tdep ? NGV * NCU : 0);
    for (int w = tid; w < W4; w += NT) {
      if (w < W1) {
        const int g = w % NGV, c = w / NGV;
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < NPV; ++a) v += sShV[g + NGV * a] * sU[a + NPV * c];
        sUg[w] = v;
      } else if (w < W2) {
        const int i = w - W1, t = i % NFG, c = i / NFG;
        const int j = t / NGF, g = t - j * NGF;
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < NPF; ++a) v += sShF[g + NGF * a] * sU[L.sPerm[a + NPF * j] + NPV * c];
        sUf[i] = v;
      } else if (w < W3) {
        const int i = w - W2, t = i % NFG, c = i / NFG;
        const int j = t / NGF, g = t - j * NGF;
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < NPF; ++a) v += sShF[g + NGF * a] * sUH[c + NCU * (a + NPF * j)];
        sUHf[i] = v;
      } else {
        const int i = w - W3, g = i % NGV, c = i / NGV;
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < NPV; ++a) v += sShV[g + NGV * a] * sSH[a + NPV * c];
        sSg[i] = v;
      }
    }
    for (int w = NT - 1 - tid; w < NGV + NFG; w += NT) {
      if (w < NGV) {
        // volume geometry: jac = det J, Xx(a, b) = jac * dxi_b/dx_a
        const int g = w;
        double Xx[ND][ND];
        const double jac = vol_geom<C>(L.ph1.sShM, sX, g, Xx);
        sJac[g] = jac;
#pragma unroll
        for (int a = 0; a < ND; ++a)
#pragma unroll
          for (int b = 0; b < ND; ++b) sXx[g * ND * ND + a * ND + b] = Xx[a][b];
        double av = 0.0;
        if (ap.av_col >= 0)
          for (int a = 0; a < NPV; ++a) av += L.ph1.sShM[g + NGV * a] * sX[a + NPV * ap.av_col];
        sAv[g] = av;
      } else {
        // face geometry: unit normal and surface Jacobian
        const int t = w - NGV;
        const int j = t / NGF, g = t - j * NGF;
        double nl[ND];
        const double jac = face_geom<C>(L.ph1.sShMF, L.sPermG, sX, j, g, nl);
#pragma unroll
        for (int i = 0; i < ND; ++i) sNl[t + NFG * i] = nl[i];
        sFJ[t] = jac;
        double av = 0.0;
        if (ap.av_col >= 0)
          for (int a = 0; a < NPF; ++a) av += L.ph1.sShMF[g + NGF * a] * sX[L.sPermG[a + NPF * j] + NPV * ap.av_col];
        sAvf[t] = av;
      }
    }
  }
  __syncthreads();

  // ---- 2b. M, C (volint.m:43-55), E (faceint.m:34-47): geometry only, so
  // they run with the physics tasks (on the wavefronts that carry the fewest
  // of them) and Rq needs no barrier of its own
  auto mce = [&]() {
    if constexpr (UQ) {
      // from the wavefront after the one with the last physics chunk
      constexpr int NCH = (NGV * NC + 63) / 64 + (NFG * NC + 63) / 64 + (NFG * NCU + 63) / 64;
      constexpr int MOFF = 64 * (NCH % (NT / 64));
      const int r0 = (tid + NT - MOFF) % NT;
      for (int idx = r0; idx < NPV * NPV; idx += NT) {
        double v = 0.0;
  #pragma unroll
        for (int g = 0; g < NGV; ++g) v += T.svv[idx + NPV * NPV * g] * sJac[g];
        v *= ap.fc_q;
        sM[idx] = v;
        if constexpr (!RO) A.M[e * NPV * NPV + idx] = v;
      }
      for (int idx = r0; idx < NPV * NPV * ND; idx += NT) {
        const int jjii = idx % (NPV * NPV), i = idx / (NPV * NPV);
        double v = 0.0;
  #pragma unroll
        for (int j = 0; j < ND; ++j)
  #pragma unroll
          for (int g = 0; g < NGV; ++g) v += T.svv[jjii + NPV * NPV * (g + NGV * (j + 1))] * sXx[g * ND * ND + i * ND + j];
        sC[idx] = v;
        if constexpr (!RO) A.C[e * NPV * NPV * ND + idx] = v;
      }
      for (int idx = r0; idx < NPV * NDF * ND; idx += NT) {
        const int va = idx % NPV;
        int rest = idx / NPV;
        const int bj = rest % NDF, i = rest / NDF;
        const int j = bj / NPF, b = bj - j * NPF;
        const int la = sLoc[va + NPV * j];
        double v = 0.0;
        if (la >= 0)
  #pragma unroll
          for (int g = 0; g < NGF; ++g) {
            const int t = j * NGF + g;
            v += sSff[la + NPF * b + NPF * NPF * g] * (sNl[t + NFG * i] * sFJ[t]);
          }
        sE[idx] = v;
        if constexpr (!RO) A.E[e * NPV * NDF * ND + idx] = v;
      }
    }
  };

  if constexpr (RO) {
    // ---- 2. physics values only: one task per volume Gauss point, then one
    // per face point
    for (int w = tid; w < NGV + NFG; w += NT) {
      if (w < NGV) {
        const int g = w;
        double u[NC], f[NCU][ND], Xx[ND][ND];
#pragma unroll
        for (int c = 0; c < NC; ++c) u[c] = sUg[g + NGV * c];
        vol_flux<C>(ap, sAv[g], u, f);
        const double jac = sJac[g];
#pragma unroll
        for (int a = 0; a < ND; ++a)
#pragma unroll
          for (int b = 0; b < ND; ++b) Xx[a][b] = sXx[g * ND * ND + a * ND + b];
#pragma unroll
        for (int c = 0; c < NCU; ++c) {
          const double sv = tdep ? sSg[g + NGV * c] - u[c] * ap.fc_u : 0.0;
          sWrk[g + NR * c] = sv * jac;
#pragma unroll
          for (int i = 0; i < ND; ++i) {
            double v = 0.0;
#pragma unroll
            for (int j = 0; j < ND; ++j) v += f[c][j] * Xx[j][i];
            sWrk[g + NGV * (i + 1) + NR * c] = v;
          }
        }
      } else {
        const int t = w - NGV;
        double nl[ND], u[NC], uh[NCU];
#pragma unroll
        for (int d = 0; d < ND; ++d) nl[d] = sNl[t + NFG * d];
#pragma unroll
        for (int c = 0; c < NC; ++c) u[c] = sUf[t + NFG * c];
#pragma unroll
        for (int c = 0; c < NCU; ++c) uh[c] = sUHf[t + NFG * c];
        D1 fa[NCU], fb[NCU];
        face_point<C>(ap, sAvf[t], nl, u, uh, sBf[t / NGF], fa, fb);
        const double jac = sFJ[t];
#pragma unroll
        for (int c = 0; c < NCU; ++c) {
          sFH[t + NFG * c] = fa[c].v * jac;
          sGH[t + NFG * c] = fb[c].v * jac;
        }
      }
    }
    mce();
  } else {
    // ---- 2. physics, one task per (Gauss point, derivative direction)
    // volume and face tasks in separate loops (their register needs do not add
    // up); task w is thread w mod NT's as in one loop over both
    // the tasks of each kind in chunks of 64, dealt to the wavefronts in turn
    // (volume chunks first, then face-UDG, then face-UH): a wavefront runs a
    // kind's code at most once a round instead of the first ones running
    // every kind
    constexpr int NWV = NT / 64;
    constexpr int CHV = (NGV * NC + 63) / 64, CHU = (NFG * NC + 63) / 64;
    const int wv = tid >> 6, ln = tid & 63;
    auto first_chunk = [&](int g0) { return ((wv - g0) % NWV + NWV) % NWV; };
    for (int ch = first_chunk(0); ch < CHV; ch += NWV) {
      const int w = 64 * ch + ln;
      if (w < NGV * NC) {
        // volume point g, direction UDG component k (volint.m:30-254)
        const int g = w % NGV, k = w / NGV;
        D1 u[NC];
  #pragma unroll
        for (int c = 0; c < NC; ++c) { u[c].v = sUg[g + NGV * c]; u[c].d[0] = (c == k) ? 1.0 : 0.0; }
        D1 f[NCU][ND];
        vol_flux<C>(ap, sAv[g], u, f);
        const double jac = sJac[g];
        double Xx[ND][ND];
  #pragma unroll
        for (int a = 0; a < ND; ++a)
  #pragma unroll
          for (int b = 0; b < ND; ++b) Xx[a][b] = sXx[g * ND * ND + a * ND + b];
  #pragma unroll
        for (int c = 0; c < NCU; ++c) {
          sWL[(c + NCU * k) + NCC * g] = (tdep && k == c) ? ap.fc_u * jac : 0.0;
  #pragma unroll
          for (int i = 0; i < ND; ++i) {
            double v = 0.0;
  #pragma unroll
            for (int j = 0; j < ND; ++j) v += f[c][j].d[0] * Xx[j][i];
            sWL[(c + NCU * k) + NCC * (g + NGV * (i + 1))] = -v;
          }
        }
        if (k == 0) {
          // zero source (euler/source.m, ns/source.m) plus the BDF time term
  #pragma unroll
          for (int c = 0; c < NCU; ++c) {
            const double s = tdep ? sSg[g + NGV * c] - u[c].v * ap.fc_u : 0.0;
            sWrk[g + NR * c] = s * jac;
  #pragma unroll
            for (int i = 0; i < ND; ++i) {
              double v = 0.0;
  #pragma unroll
              for (int j = 0; j < ND; ++j) v += f[c][j].v * Xx[j][i];
              sWrk[g + NGV * (i + 1) + NR * c] = v;
            }
          }
        }
      }
    }
    // face point t, direction k: UDG component k (k < NC, uh constant; tasks
    // NGV*NC .. +NFG*NC) or UH component k - NC (u constant; the rest), each
    // kind in a loop of its own (faceint.m:55-277)
    auto face_out = [&](int t, int k, const D1* fa, const D1* fb) {
      const double jac = sFJ[t];
  #pragma unroll
      for (int c = 0; c < NCU; ++c) {
        if (k < NC) {
          sWL[(c + NCU * k) + NCC * (NR + t)] = fa[c].d[0] * jac;
          sGHu[t + NFG * (c + NCU * k)] = fb[c].d[0] * jac;
        } else {
          sFHh[t + NFG * (c + NCU * (k - NC))] = fa[c].d[0] * jac;
          sGHh[t + NFG * (c + NCU * (k - NC))] = fb[c].d[0] * jac;
        }
        if (k == 0) {
          sFH[t + NFG * c] = fa[c].v * jac;
          sGH[t + NFG * c] = fb[c].v * jac;
        }
      }
    };
    for (int ch = first_chunk(CHV); ch < CHU; ch += NWV) {
      const int i = 64 * ch + ln, t = i % NFG, k = i / NFG;
      if (i >= NFG * NC) continue;
      double nl[ND];
  #pragma unroll
      for (int d = 0; d < ND; ++d) nl[d] = sNl[t + NFG * d];
      D1 u[NC], fa[NCU], fb[NCU];
      double uh[NCU];
  #pragma unroll
      for (int c = 0; c < NC; ++c) { u[c].v = sUf[t + NFG * c]; u[c].d[0] = (c == k) ? 1.0 : 0.0; }
  #pragma unroll
      for (int c = 0; c < NCU; ++c) uh[c] = sUHf[t + NFG * c];
      face_point<C>(ap, sAvf[t], nl, u, uh, sBf[t / NGF], fa, fb);
      face_out(t, k, fa, fb);
    }
    for (int ch = first_chunk(CHV + CHU); ch < (NFG * NCU + 63) / 64; ch += NWV) {
      const int i = NFG * NC + 64 * ch + ln, t = i % NFG, k = i / NFG;
      if (i >= NFG * NDIRF) continue;
      double nl[ND];
  #pragma unroll
      for (int d = 0; d < ND; ++d) nl[d] = sNl[t + NFG * d];
      double u[NC];
      D1 uh[NCU], fa[NCU], fb[NCU];
  #pragma unroll
      for (int c = 0; c < NC; ++c) u[c] = sUf[t + NFG * c];
  #pragma unroll
      for (int c = 0; c < NCU; ++c) { uh[c].v = sUHf[t + NFG * c]; uh[c].d[0] = (c + NC == k) ? 1.0 : 0.0; }
      face_point<C>(ap, sAvf[t], nl, u, uh, sBf[t / NGF], fa, fb);
      face_out(t, k, fa, fb);
    }
    mce();
  }
  __syncthreads();

  // ---- 3b. Ru = shapvg*wrk - sum_faces shapfg*FH (volint.m:258, faceint.m:302-330)
  double* Ru = A.Ru + e * NU;
  for (int idx = tid; idx < NU; idx += NT) {
    const int a = idx % NPV, c = idx / NPV;
    double v = 0.0;
#pragma unroll
    for (int r = 0; r < NR; ++r) v += T.shapvg[a + NPV * r] * sWrk[r + NR * c];
#pragma unroll
    for (int j = 0; j < NFE; ++j) {
      const int la = sLoc[a + NPV * j];
      if (la >= 0)
#pragma unroll
        for (int g = 0; g < NGF; ++g) v -= T.shapfg[la + NPF * g] * sFH[j * NGF + g + NFG * c];
    }
    Ru[idx] = v;
  }
  // ---- 3c. Rh = -shapfg * (fbou | fhat) (faceint.m:335-339)
  double* Rh = A.Rh + e * NH;
  for (int idx = (tid + NT - 64 % NT) % NT; idx < NH; idx += NT) {     // from the second wavefront on (Ru takes the first)
    const int c = idx % NCU, ab = idx / NCU, j = ab / NPF, a = ab - j * NPF;
    double v = 0.0;
#pragma unroll
    for (int g = 0; g < NGF; ++g) v += T.shapfg[a + NPF * g] * sGH[j * NGF + g + NFG * c];
    Rh[idx] = RO ? v : -v;
  }
  if constexpr (UQ) {
    // ---- 3h. Rq = M (SH_q/fc_q - q) + C u - E uh (volint.m:57-60, faceint.m:48-51), from the
    // third wavefront on
    for (int idx = (tid + NT - 128 % NT) % NT; idx < NPV * NCU * ND; idx += NT) {
      const int a = idx % NPV;
      int rest = idx / NPV;
      const int c = rest % NCU, i = rest / NCU;
      const int qc = NCU + c + NCU * i;
      double v = 0.0, w = 0.0, z = 0.0;
      for (int b = 0; b < NPV; ++b) v += sM[a + NPV * b] * (sSH[b + NPV * qc] / ap.fc_q - sU[b + NPV * qc]);
      for (int b = 0; b < NPV; ++b) w += sC[a + NPV * b + NPV * NPV * i] * sU[b + NPV * c];
      for (int b = 0; b < NDF; ++b) z += sE[a + NPV * b + NPV * NDF * i] * sUH[c + NCU * b];
      A.Rq[e * NPV * NCU * ND + idx] = v + w - z;
    }
  }
  if constexpr (!RO) {
    // ---- 3a. BD (npv,ncu,npv,nc) = volume + face blocks (volint.m:262,
    // faceint.m:316,326) as one FP64 MFMA product per element:
    //   BD(jj, c, ii, k) = sum_r' WL(r', (c, k)) SV(r', (jj, ii)),
    // r' over the volume rows r < NR (WL = wrl, SV = shapvgdotshapvl) and the
    // face points NR + t (WL = the fhat Jacobian, SV = shapfgdotshapfc of
    // (jj, ii) where both lie on the point's face, else 0).  SV is the master's
    // B-fragment table (T.bdv); WL^T's A fragments come from LDS.  Each
    // wavefront takes row tiles R of (jj, ii) and keeps their SV fragments in
    // registers across the column tiles of (c, k).
    // the gather keys of F, GK and H (FghKeys): F's loaded before the BD
    // products (whose MFMA hides their latency), GK's and H's one array ahead
    constexpr bool FBK = FbShape<C>::USE;
    constexpr bool STG = TILE && FaceStage<C>::USE;
    using KT_ = FghKeys<C, FBK, STG>;
    constexpr int IF = TILE ? (KT_::NF + NT - 1) / NT : 1, IG = TILE ? (KT_::NG + NT - 1) / NT : 1;
    constexpr int IH = TILE ? (KT_::NHK + NT - 1) / NT : 1;
    auto load_keys = [&](const int32_t* key, int nv, int* k, int n) {
  #pragma unroll
      for (int it = 0; it < n; ++it) k[it] = tid + it * NT < nv ? key[tid + it * NT] : -1;
    };
    // in registers when they fit (at most 16 a thread), else loaded in the
    // loops; the staged path (STG) loads each copy loop's keys itself, ahead
    // of its staging products
    constexpr bool KPRE = IF + IG + IH <= 16;
    int kf[KPRE ? IF : 1], kg[KPRE ? IG : 1], kh[KPRE ? IH : 1];
    if constexpr (TILE && KPRE && !STG) load_keys(FghKeyTab<C, FBK, STG>::tab.f, KT_::NF, kf, IF);
    using BS = BdShape<C>;
    constexpr int KSB = BS::KSB, KT = BS::KT, NVV = BS::NVV;
    double* BD = A.BD + e * (int64_t)NU * C::NCOL;
    const int wv = tid >> 6, ln = tid & 63, gq = ln >> 4, n = ln & 15;
    for (int R = wv; R < BS::RTB; R += NT / 64) {
      double b[KSB];
  #pragma unroll
      for (int s = 0; s < KSB; ++s) b[s] = T.bdv[(R * KSB + s) * 64 + ln];
      const int col = 16 * R + n, jj = col % NPV, ii = col / NPV;
  #pragma unroll
      for (int Cc = 0; Cc < BS::CTB; ++Cc) {
        d4 acc = {0.0, 0.0, 0.0, 0.0};
  #pragma unroll
        for (int s = 0; s < KSB; ++s) {
          const int r = 4 * s + gq;
          const double a = r < KT ? sWL[16 * Cc + n + NCC * r] : 0.0;   // rows past NCC only feed unused rows
          acc = mfma64(a, b[s], acc);
        }
  #pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int cc = 16 * Cc + 4 * i + gq, c = cc % NCU, k = cc / NCU;
          if (cc < NCC && col < NVV) BD[jj + NPV * (c + NCU * (ii + NPV * k))] = acc[i];
        }
      }
    }
    if constexpr (TILE && KPRE && !STG) load_keys(FghKeyTab<C, FBK, STG>::tab.gk, KT_::NG, kg, IG);
    if constexpr (STG) {
      // F, GK and H from the staged face products (FaceStage): one LDS read
      // an entry
      using FS = FaceStage<C>;
      using S = S64Shape<C>;
      constexpr int PT = FS::PT;
      double* const sP = sWL;
      // P_f over the NX columns of X into dst; tasks (face, row tile, column
      // tile) dealt over the wavefronts
      auto stage = [&](auto nxc, const double* X, double* dst) {
        constexpr int NX = decltype(nxc)::value, CB = (NX + 15) / 16, NTASK = NFE * FS::RTP * CB;
        for (int task = wv; task < NTASK; task += NT / 64) {
          const int cb = task % CB, rt = (task / CB) % FS::RTP, f = task / (CB * FS::RTP);
          const int pr = 16 * rt + n, col = 16 * cb + n;
          d4 acc = {0.0, 0.0, 0.0, 0.0};
  #pragma unroll
          for (int s = 0; s < FS::KSF; ++s) {
            const int k = 4 * s + gq;
            const double a = (pr < PT && k < NGF) ? sSff[pr + PT * k] : 0.0;
            const double b = (col < NX && k < NGF) ? X[f * NGF + k + NFG * col] : 0.0;
            acc = mfma64(a, b, acc);
          }
  #pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int pp = 16 * rt + 4 * i + gq;
            if (pp < PT && col < NX) dst[(f * PT + pp) * NX + col] = acc[i];
          }
        }
      };
      // each copy loop's keys, loaded ahead of its staging products and
      // barriers, which hide their latency
      int kgs[IG];
      load_keys(FghKeyTab<C, FBK, STG>::tab.gk, KT_::NG, kgs, IG);
      __syncthreads();                        // BD's reads of sWL done
      stage(std::integral_constant<int, FS::NXG>{}, sGHu, sP);
      __syncthreads();
      {
        // GK: A fragments of the AE product over the face-sparse columns
        constexpr int NV = S::KSP * S::TRH * 64;
        const unsigned char* tab = L.sAek;
        double* GK = A.GK + e * (int64_t)NV;
  #pragma unroll
        for (int it = 0; it < IG; ++it) {
          const int idx = tid + it * NT;
          const int kv = kgs[it];
          if (idx >= NV) break;
          const int kcol = kv >= 0 ? tab[kv & 1023] : 255;
          double v = 0.0;
          if (kcol != 255) {
            const int j = (kv >> 10) & 15, b = (kv >> 14) & 15, c = kv >> 18;
            const int va = kcol % NPV, k = kcol / NPV;
            const int la = sLoc[va + NPV * j];
            if (la >= 0) v = sP[(j * PT + la + NPF * b) * FS::NXG + c + NCU * k];
          }
          GK[idx] = v;
        }
      }
      int kfs[IF], khs[IH];
      load_keys(FghKeyTab<C, FBK, STG>::tab.f, KT_::NF, kfs, IF);
      load_keys(FghKeyTab<C, FBK, STG>::tab.h, KT_::NHK, khs, IH);
      __syncthreads();                        // GK's products read
      stage(std::integral_constant<int, FS::NXF>{}, sFHh, sP);
      stage(std::integral_constant<int, FS::NXH>{}, sGHh, sP + NFE * PT * FS::NXF);
      __syncthreads();
      {
        // F's face blocks
        constexpr int NV = KT_::NF;
        double* F = A.F + e * (int64_t)NV;
  #pragma unroll
        for (int it = 0; it < IF; ++it) {
          const int idx = tid + it * NT;
          const int kv = kfs[it];
          if (idx < NV) F[idx] = sP[kv];
        }
      }
      {
        // H's diagonal face blocks
        constexpr int NV = KT_::NHK;
        const double* sPH = sP + NFE * PT * FS::NXF;
        double* H = A.H + e * (int64_t)NV;
  #pragma unroll
        for (int it = 0; it < IH; ++it) {
          const int idx = tid + it * NT;
          const int kv = khs[it];
          if (idx < NV) H[idx] = kv >= 0 ? sPH[kv] : 0.0;
        }
      }
    }
    // ---- 3d. F (npv,ncu,ncu,ndf) (faceint.m:320,327); F(va, c, k, bj) is u row
    // va + NPV c, trace column k + NCU bj
    auto f_entry = [&](int va, int c, int k, int bj) {
      const int j = bj / NPF, b = bj - j * NPF;
      const int la = sLoc[va + NPV * j];
      double v = 0.0;
      if (la >= 0)
  #pragma unroll
        for (int g = 0; g < NGF; ++g) v += sSff[la + NPF * b + NPF * NPF * g] * sFHh[j * NGF + g + NFG * (c + NCU * k)];
      return v;
    };
    // a keyed entry: sum_g shapfgdotshapfc(so + NPF^2 g) X(xo + g) (FghKeys)
    auto keyed = [&](int so, const double* X) {
      double v = 0.0;
  #pragma unroll
      for (int g = 0; g < NGF; ++g) v += sSff[so + NPF * NPF * g] * X[g];
      return v;
    };
    if constexpr (STG) {
    } else if constexpr (TILE) {
      // face blocks: F_f(q, r) = F(u row of q on face f, trace column f NB + r)
      constexpr int NV = KT_::NF;
      double* F = A.F + e * (int64_t)NV;
      if constexpr (KPRE) load_keys(FghKeyTab<C, FBK, STG>::tab.h, KT_::NHK, kh, IH);
  #pragma unroll
      for (int it = 0; it < IF; ++it) {
        const int idx = tid + it * NT;
        const int kv = KPRE ? kf[KPRE ? it : 0] : (idx < NV ? FghKeyTab<C, FBK, STG>::tab.f[idx] : -1);
        if (idx < NV) F[idx] = keyed(kv & 255, sFHh + (kv >> 8));
      }
    } else {
      double* F = A.F + e * (int64_t)NU * NH;
      for (int idx = tid; idx < NU * NH; idx += NT) {
        const int va = idx % NPV;
        int rest = idx / NPV;
        const int c = rest % NCU;
        rest /= NCU;
        const int k = rest % NCU;
        F[idx] = f_entry(va, c, k, rest / NCU);
      }
    }
    // ---- 3e. GK (ncu,ndf,npv,nc) (faceint.m:342-358); GK(c, bj, va, k) is row
    // c + NCU bj, column va + NPV k
    auto gk_entry = [&](int c, int bj, int va, int k) {
      const int j = bj / NPF, b = bj - j * NPF;
      const int la = sLoc[va + NPV * j];
      double v = 0.0;
      if (la >= 0)
  #pragma unroll
        for (int g = 0; g < NGF; ++g) v += sSff[la + NPF * b + NPF * NPF * g] * sGHu[j * NGF + g + NFG * (c + NCU * k)];
      return v;
    };
    if constexpr (STG) {
    } else if constexpr (TILE && FbShape<C>::USE) {
      // face blocks: GK_f(r, q) = GK(trace row f NB + r, u column of q on face f)
      constexpr int NV = (int)FbShape<C>::BLK;
      double* GK = A.GK + e * (int64_t)NV;
  #pragma unroll
      for (int it = 0; it < IG; ++it) {
        const int idx = tid + it * NT;
        const int kv = KPRE ? kg[KPRE ? it : 0] : (idx < NV ? FghKeyTab<C, FBK, STG>::tab.gk[idx] : -1);
        if (idx < NV) GK[idx] = keyed(kv & 255, sGHu + (kv >> 8));
      }
    } else if constexpr (TILE) {
      // A fragments of the AE product over the face-sparse columns (the
      // master's AE k-step table gives each entry's column (va, k))
      using S = S64Shape<C>;
      constexpr int NV = S::KSP * S::TRH * 64;
      const unsigned char* tab = L.sAek;
      double* GK = A.GK + e * (int64_t)NV;
  #pragma unroll
      for (int it = 0; it < IG; ++it) {
        const int idx = tid + it * NT;
        const int kv = KPRE ? kg[KPRE ? it : 0] : (idx < NV ? FghKeyTab<C, FBK, STG>::tab.gk[idx] : -1);
        if (idx >= NV) break;
        const int kcol = kv >= 0 ? tab[kv & 1023] : 255;
        double v = 0.0;
        if (kcol != 255) {
          const int j = (kv >> 10) & 15, b = (kv >> 14) & 15, c = kv >> 18;
          const int va = kcol % NPV, k = kcol / NPV;
          const int la = sLoc[va + NPV * j];
          if (la >= 0) v = keyed(la + NPF * b, sGHu + j * NGF + NFG * (c + NCU * k));
        }
        GK[idx] = v;
      }
    } else {
      double* GK = A.GK + e * (int64_t)NH * C::NCOL;
      for (int idx = tid; idx < NH * C::NCOL; idx += NT) {
        const int c = idx % NCU;
        int rest = idx / NCU;
        const int bj = rest % NDF;
        rest /= NDF;
        GK[idx] = gk_entry(c, bj, rest % NPV, rest / NPV);
      }
    }
    // ---- 3f. H (ncu,ndf,ncu,ndf), block diagonal per face (faceint.m:347-359);
    // H(c, aj, k, bj) is row c + NCU aj, column k + NCU bj
    auto h_entry = [&](int c, int aj, int k, int bj) {
      const int j = aj / NPF, a = aj - j * NPF, j2 = bj / NPF, b = bj - j2 * NPF;
      double v = 0.0;
      if (j == j2)
  #pragma unroll
        for (int g = 0; g < NGF; ++g) v += sSff[a + NPF * b + NPF * NPF * g] * sGHh[j * NGF + g + NFG * (c + NCU * k)];
      return v;
    };
    if constexpr (STG) {
    } else if constexpr (TILE) {
      // face blocks H_f(r, r') = H(f NB + r, f NB + r')
      constexpr int NV = KT_::NHK;
      double* H = A.H + e * (int64_t)NV;
  #pragma unroll
      for (int it = 0; it < IH; ++it) {
        const int idx = tid + it * NT;
        const int kv = KPRE ? kh[KPRE ? it : 0] : (idx < NV ? FghKeyTab<C, FBK, STG>::tab.h[idx] : -1);
        if (idx < NV) H[idx] = kv >= 0 ? keyed(kv & 255, sGHh + (kv >> 8)) : 0.0;
      }
    } else {
      double* H = A.H + e * (int64_t)NH * NH;
      for (int idx = tid; idx < NH * NH; idx += NT) {
        const int c = idx % NCU;
        int rest = idx / NCU;
        const int aj = rest % NDF;
        rest /= NDF;
        H[idx] = h_entry(c, aj, rest % NCU, rest / NCU);
      }
    }
  }
}

template <class C, bool TILE = false>
__global__ __launch_bounds__(QuadThreads<C>::N) __attribute__((amdgpu_waves_per_eu(QuadWpe<C>::N1))) void quad_kernel(QuadArgs A, DevMaster T, DevApp ap) {
  __shared__ QuadLds<C> lds;
  quad_body<C, false, TILE, QuadThreads<C>::N>(A, T, ap, blockIdx.x, lds);
}

// residual only: Ru, Rh (hdg_residual.m's sign), Rq
template <class C>
__global__ __launch_bounds__(QuadThreads<C>::N) __attribute__((amdgpu_waves_per_eu(5))) void resid_kernel(QuadArgs A, DevMaster T, DevApp ap) {
  __shared__ QuadLds<C, true> lds;
  quad_body<C, true, false, QuadThreads<C>::N>(A, T, ap, blockIdx.x, lds);
}

}  // namespace hdgx
