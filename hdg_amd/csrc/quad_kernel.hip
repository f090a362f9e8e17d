// Element matrices of the HDG elemental path on gfx950: one workgroup per
// element computes everything volint + faceint produce (kernel/volint.m:1-263,
// kernel/faceint.m:1-360) straight into the layouts the Schur step consumes.
//
// Dataflow per element:
//   1. stage dgnodes, UDG, the gathered UH (and SH) of the element in LDS;
//   2. one thread per volume Gauss point: geometry (volgeom), u at the point,
//      flux + Jacobian, then the weighted rows wrk/wrl (volint.m:241-254);
//      concurrently one thread per face Gauss point (another wavefront):
//      facegeom, u and uh traces, fhat (all faces) and fbou (boundary faces),
//      pre-multiplied by the surface Jacobian;
//   3. every output entry is a gather-sum over the staged rows, written once
//      with consecutive threads on consecutive addresses.  The face scatter of
//      faceint.m:322-360 (perm) becomes a per-node lookup loc(v, j) of node v
//      on face j, so no entry is read-modify-written.
#include <hip/hip_runtime.h>

#include "hdgx_device.h"
#include "hdgx_kernels.h"

namespace hdgx {

constexpr int QUAD_THREADS = 256;
constexpr int QUAD_FACE_T0 = 128;   // face-point threads start in wavefront 2

// f (NCU x ND) and f_udg (NCU x ND x NC) at one point (euler/flux.m, ns/flux2d.m,
// laplacian AV f += eps_av * q).
template <class C>
HDGX_DEV void flux_point(const DevApp& ap, double av, const double* u, double f[C::NCU][C::ND],
                         double fu[C::NCU][C::ND][C::NC]) {
  constexpr int ND = C::ND, NCU = C::NCU, NC = C::NC;
  if constexpr (C::PHYS == PHYS_EULER) {
    Dual<NCU> w[NCU];
#pragma unroll
    for (int c = 0; c < NCU; ++c) w[c] = Dual<NCU>::seed(u[c], c);
    Dual<NCU> fl[NCU][ND];
    euler_flux<ND>(ap.gam, w, fl);
#pragma unroll
    for (int c = 0; c < NCU; ++c)
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        f[c][d] = fl[c][d].v;
#pragma unroll
        for (int k = 0; k < NC; ++k) fu[c][d][k] = fl[c][d].d[k];
      }
  } else {
    // pass 1: derivatives w.r.t. u (q held fixed)
    {
      Dual<4> w[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) w[c] = Dual<4>::seed(u[c], c);
      Dual<4> fl[4][2];
      ns2d_flux(ap.gam, ap.Re, ap.Pr, w, u + 4, fl);
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int d = 0; d < 2; ++d) {
          f[c][d] = fl[c][d].v + av * u[4 + c + 4 * d];
#pragma unroll
          for (int k = 0; k < 4; ++k) fu[c][d][k] = fl[c][d].d[k];
        }
    }
    // pass 2: derivatives w.r.t. q (linear in q)
    {
      Dual<8> w[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) w[c] = Dual<8>::seed(u[4 + c], c);
      Dual<8> fl[4][2];
      ns2d_flux(ap.gam, ap.Re, ap.Pr, u, w, fl);
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int d = 0; d < 2; ++d)
#pragma unroll
          for (int k = 0; k < 8; ++k) fu[c][d][4 + k] = fl[c][d].d[k] + ((k == c + 4 * d) ? av : 0.0);
    }
  }
}

// fhat = f(uh[,q]).n + tau (u - uh) (euler/fhat.m caseFlux 1/caseStab 4, ns/fhat.m:49-67)
template <class C>
HDGX_DEV void fhat_point(const DevApp& ap, double av, const double* n, const double* u, const double* uh,
                         double fh[C::NCU], double fhu[C::NCU][C::NC], double fhh[C::NCU][C::NCU]) {
  constexpr int ND = C::ND, NCU = C::NCU, NC = C::NC;
  const double tau = ap.tau;
#pragma unroll
  for (int c = 0; c < NCU; ++c)
#pragma unroll
    for (int k = 0; k < NC; ++k) fhu[c][k] = (k == c) ? tau : 0.0;
  if constexpr (C::PHYS == PHYS_EULER) {
    Dual<NCU> w[NCU];
#pragma unroll
    for (int c = 0; c < NCU; ++c) w[c] = Dual<NCU>::seed(uh[c], c);
    Dual<NCU> fl[NCU][ND];
    euler_flux<ND>(ap.gam, w, fl);
#pragma unroll
    for (int c = 0; c < NCU; ++c) {
      Dual<NCU> fn(0.0);
#pragma unroll
      for (int d = 0; d < ND; ++d) fn = fn + fl[c][d] * n[d];
      fh[c] = fn.v + tau * (u[c] - uh[c]);
#pragma unroll
      for (int k = 0; k < NCU; ++k) fhh[c][k] = fn.d[k] - ((c == k) ? tau : 0.0);
    }
  } else {
    {
      Dual<4> w[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) w[c] = Dual<4>::seed(uh[c], c);
      Dual<4> fl[4][2];
      ns2d_flux(ap.gam, ap.Re, ap.Pr, w, u + 4, fl);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        Dual<4> fn = fl[c][0] * n[0] + fl[c][1] * n[1];
        fh[c] = fn.v + av * (u[4 + c] * n[0] + u[8 + c] * n[1]) + tau * (u[c] - uh[c]);
#pragma unroll
        for (int k = 0; k < 4; ++k) fhh[c][k] = fn.d[k] - ((c == k) ? tau : 0.0);
      }
    }
    {
      Dual<8> w[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) w[c] = Dual<8>::seed(u[4 + c], c);
      Dual<8> fl[4][2];
      ns2d_flux(ap.gam, ap.Re, ap.Pr, uh, w, fl);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        Dual<8> fn = fl[c][0] * n[0] + fl[c][1] * n[1];
#pragma unroll
        for (int k = 0; k < 8; ++k)
          fhu[c][4 + k] = fn.d[k] + ((k == c) ? av * n[0] : 0.0) + ((k == c + 4) ? av * n[1] : 0.0);
      }
    }
  }
}

// Boundary flux: far field (ib 1: euler/fbou.m:24-29 LLF in 2D Euler, Roe form
// for NS and 3D Euler: ns/fbou.m:37-46, fbou_euler3d.c:20-60) and the 2D Euler
// inviscid wall (ib 2: euler/fbou.m:30-48).  Host validation rejects others.
template <class C>
HDGX_DEV void fbou_point(const DevApp& ap, int ib, const double* uinf, const double* n, const double* u,
                         const double* uh, double fh[C::NCU], double fhu[C::NCU][C::NC],
                         double fhh[C::NCU][C::NCU]) {
  constexpr int ND = C::ND, NCU = C::NCU, NC = C::NC;
#pragma unroll
  for (int c = 0; c < NCU; ++c)
#pragma unroll
    for (int k = 0; k < NC; ++k) fhu[c][k] = 0.0;
  if (ib == 2) {
    if constexpr (C::PHYS == PHYS_EULER && ND == 2) {
      double un = u[1] * n[0] + u[2] * n[1];
      double ub[4] = {u[0], u[1] - n[0] * un, u[2] - n[1] * un, u[3]};
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        fh[c] = ub[c] - uh[c];
#pragma unroll
        for (int k = 0; k < 4; ++k) fhh[c][k] = (c == k) ? -1.0 : 0.0;
      }
      fhu[0][0] = 1.0;
      fhu[1][1] = 1.0 - n[0] * n[0];
      fhu[1][2] = -n[0] * n[1];
      fhu[2][1] = -n[0] * n[1];
      fhu[2][2] = 1.0 - n[1] * n[1];
      fhu[3][3] = 1.0;
    }
    return;
  }
  Dual<NCU> w[NCU];
#pragma unroll
  for (int c = 0; c < NCU; ++c) w[c] = Dual<NCU>::seed(uh[c], c);
  Dual<NCU> an[NCU][NCU], An[NCU][NCU];
  if constexpr (C::PHYS == PHYS_EULER && ND == 2) {
    const double gam1 = ap.gam - 1.0;
    Dual<NCU> r1 = 1.0 / w[0];
    Dual<NCU> uv = w[1] * r1, vv = w[2] * r1;
    Dual<NCU> af = 0.5 * (uv * uv + vv * vv);
    Dual<NCU> p = gam1 * (w[3] - w[0] * af);
    Dual<NCU> c = dsqrt(ap.gam * p * r1);
    Dual<NCU> un = uv * n[0] + vv * n[1];
    Dual<NCU> l2 = dabs(un) + c, l3 = un - c;
#pragma unroll
    for (int i = 0; i < NCU; ++i)
#pragma unroll
      for (int k = 0; k < NCU; ++k) { an[i][k] = (i == k) ? l2 : Dual<NCU>(0.0); An[i][k] = (i == k) ? l3 : Dual<NCU>(0.0); }
  } else {
    roe_an<ND>(ap.gam, ap.eps, 0, w, n, an);
    roe_an<ND>(ap.gam, ap.eps, 1, w, n, An);
  }
#pragma unroll
  for (int i = 0; i < NCU; ++i) {
    Dual<NCU> s(0.0);
#pragma unroll
    for (int k = 0; k < NCU; ++k) s = s + (an[i][k] + An[i][k]) * (u[k] - w[k]) - (an[i][k] - An[i][k]) * (uinf[k] - w[k]);
    fh[i] = s.v;
#pragma unroll
    for (int k = 0; k < NCU; ++k) { fhu[i][k] = an[i][k].v + An[i][k].v; fhh[i][k] = s.d[k]; }
  }
}

template <class C>
__global__ __launch_bounds__(QUAD_THREADS) void quad_kernel(QuadArgs A, DevMaster T, DevApp ap) {
  constexpr int ND = C::ND, NPV = C::NPV, NPF = C::NPF, NFE = C::NFE, NDF = C::NDF;
  constexpr int NGV = C::NGV, NGF = C::NGF, NCU = C::NCU, NC = C::NC, NU = C::NU, NH = C::NH;
  constexpr int NR = C::NR, NFG = NFE * NGF, NT = QUAD_THREADS;
  constexpr bool UQ = C::UQ;
  static_assert(NGV <= QUAD_FACE_T0 && NFG <= NT - QUAD_FACE_T0, "too many quadrature points");

  const int tid = threadIdx.x;
  const int64_t e = blockIdx.x;
  if (e >= A.ne) return;
  const int ncd = ap.ncd;

  __shared__ double sX[NPV * (ND + 1)];
  __shared__ double sU[NPV * NC];
  __shared__ double sUH[NCU * NDF];
  __shared__ double sSH[NPV * NC];
  __shared__ double sJac[NGV];
  __shared__ double sXx[NGV * ND * ND];
  __shared__ double sWrk[NR * NCU];
  __shared__ double sWrl[NR * NCU * NC];
  __shared__ double sNJ[NFG * ND];
  __shared__ double sFH[NFG * NCU], sFHu[NFG * NCU * NC], sFHh[NFG * NCU * NCU];
  __shared__ double sGH[NFG * NCU], sGHu[NFG * NCU * NC], sGHh[NFG * NCU * NCU];
  __shared__ double sM[UQ ? NPV * NPV : 1], sC[UQ ? NPV * NPV * ND : 1], sE[UQ ? NPV * NDF * ND : 1];
  __shared__ int32_t sBf[NFE];

  // ---- 1. stage the element's inputs
  for (int i = tid; i < NPV * ncd; i += NT) sX[i] = A.dgnodes[e * NPV * ncd + i];
  for (int i = tid; i < NPV * NC; i += NT) sU[i] = A.udg[e * NPV * NC + i];
  for (int i = tid; i < NCU * NDF; i += NT) sUH[i] = A.uh[e * NH + i];
  if (UQ || ap.tdep)
    for (int i = tid; i < NPV * NC; i += NT) sSH[i] = A.sh ? A.sh[e * NPV * NC + i] : 0.0;
  if (tid < NFE) sBf[tid] = A.bf[e * NFE + tid];
  __syncthreads();

  // ---- 2a. volume Gauss points (volgeom.m:22-59, volint.m:30-254)
  if (tid < NGV) {
    const int g = tid;
    double J[ND][ND];
#pragma unroll
    for (int j = 0; j < ND; ++j)
#pragma unroll
      for (int i = 0; i < ND; ++i) {
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < NPV; ++a) v += T.shapmv[g + NGV * a + NGV * NPV * (1 + j)] * sX[a + NPV * i];
        J[j][i] = v;   // dx_i / dxi_j
      }
    double jac, Xx[ND][ND];   // Xx[a][b] = jac * dxi_b / dx_a
    if constexpr (ND == 2) {
      jac = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      Xx[0][0] = J[1][1]; Xx[1][0] = -J[1][0]; Xx[0][1] = -J[0][1]; Xx[1][1] = J[0][0];
    } else {
      jac = J[0][0] * J[1][1] * J[2][2] - J[0][0] * J[2][1] * J[1][2] + J[1][0] * J[2][1] * J[0][2] -
            J[1][0] * J[0][1] * J[2][2] + J[2][0] * J[0][1] * J[1][2] - J[2][0] * J[1][1] * J[0][2];
      Xx[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      Xx[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      Xx[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      Xx[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
      Xx[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
      Xx[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
      Xx[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
      Xx[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
      Xx[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    }
    sJac[g] = jac;
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int b = 0; b < ND; ++b) sXx[g * ND * ND + a * ND + b] = Xx[a][b];
    double av = 0.0;
    if (ap.av_col >= 0) {
      for (int a = 0; a < NPV; ++a) av += T.shapmv[g + NGV * a] * sX[a + NPV * ap.av_col];
    }
    double ug[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < NPV; ++a) v += T.shapvt[g + NGV * a] * sU[a + NPV * c];
      ug[c] = v;
    }
    double f[NCU][ND], fu[NCU][ND][NC];
    flux_point<C>(ap, av, ug, f, fu);
    // zero source (euler/source.m, ns/source.m) plus the BDF time term
    for (int c = 0; c < NCU; ++c) {
      double s = 0.0;
      if (ap.tdep) {
        double st = 0.0;
        for (int a = 0; a < NPV; ++a) st += T.shapvt[g + NGV * a] * sSH[a + NPV * c];
        s = st - ug[c] * ap.fc_u;
      }
      sWrk[g + NR * c] = s * jac;
      for (int k = 0; k < NC; ++k) sWrl[g + NR * (c + NCU * k)] = (ap.tdep && k == c) ? ap.fc_u * jac : 0.0;
    }
#pragma unroll
    for (int i = 0; i < ND; ++i)
#pragma unroll
      for (int c = 0; c < NCU; ++c) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < ND; ++j) v += f[c][j] * Xx[j][i];
        sWrk[g + NGV * (i + 1) + NR * c] = v;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          double w = 0.0;
#pragma unroll
          for (int j = 0; j < ND; ++j) w += fu[c][j][k] * Xx[j][i];
          sWrl[g + NGV * (i + 1) + NR * (c + NCU * k)] = -w;
        }
      }
  }

  // ---- 2b. face Gauss points (facegeom.m:13-52, faceint.m:55-277)
  if (tid >= QUAD_FACE_T0 && tid < QUAD_FACE_T0 + NFG) {
    const int t = tid - QUAD_FACE_T0;
    const int j = t / NGF, g = t - j * NGF;
    double dpg[ND][ND > 1 ? ND - 1 : 1];
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
      for (int k = 0; k < ND - 1; ++k) {
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < NPF; ++a) v += T.shapmf[g + NGF * a + NGF * NPF * (1 + k)] * sX[T.permgeom[a + NPF * j] + NPV * d];
        dpg[d][k] = v;
      }
    double nl[ND], jac;
    if constexpr (ND == 2) {
      jac = sqrt(dpg[0][0] * dpg[0][0] + dpg[1][0] * dpg[1][0]);
      nl[0] = dpg[1][0] / jac;
      nl[1] = -dpg[0][0] / jac;
    } else {
      double n0 = dpg[1][0] * dpg[2][1] - dpg[2][0] * dpg[1][1];
      double n1 = dpg[2][0] * dpg[0][1] - dpg[0][0] * dpg[2][1];
      double n2 = dpg[0][0] * dpg[1][1] - dpg[1][0] * dpg[0][1];
      jac = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      nl[0] = n0 / jac; nl[1] = n1 / jac; nl[2] = n2 / jac;
    }
#pragma unroll
    for (int i = 0; i < ND; ++i) sNJ[t + NFG * i] = nl[i] * jac;
    double av = 0.0;
    if (ap.av_col >= 0)
      for (int a = 0; a < NPF; ++a) av += T.shapmf[g + NGF * a] * sX[T.permgeom[a + NPF * j] + NPV * ap.av_col];
    double ug[NC], uhg[NCU];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < NPF; ++a) v += T.shapft[g + NGF * a] * sU[T.perm[a + NPF * j] + NPV * c];
      ug[c] = v;
    }
#pragma unroll
    for (int c = 0; c < NCU; ++c) {
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < NPF; ++a) v += T.shapft[g + NGF * a] * sUH[c + NCU * (a + NPF * j)];
      uhg[c] = v;
    }
    double fh[NCU], fhu[NCU][NC], fhh[NCU][NCU];
    fhat_point<C>(ap, av, nl, ug, uhg, fh, fhu, fhh);
#pragma unroll
    for (int c = 0; c < NCU; ++c) {
      sFH[t + NFG * c] = fh[c] * jac;
#pragma unroll
      for (int k = 0; k < NC; ++k) sFHu[t + NFG * (c + NCU * k)] = fhu[c][k] * jac;
#pragma unroll
      for (int k = 0; k < NCU; ++k) sFHh[t + NFG * (c + NCU * k)] = fhh[c][k] * jac;
    }
    const int bfj = sBf[j];
    if (bfj < 0) {
      const int id = -bfj - 1;
      const int ib = ap.bcm[id];
      double uinf[NCU];
#pragma unroll
      for (int c = 0; c < NCU; ++c) uinf[c] = ap.bcs[id * 5 + c];
      fbou_point<C>(ap, ib, uinf, nl, ug, uhg, fh, fhu, fhh);
    }
#pragma unroll
    for (int c = 0; c < NCU; ++c) {
      sGH[t + NFG * c] = fh[c] * jac;
#pragma unroll
      for (int k = 0; k < NC; ++k) sGHu[t + NFG * (c + NCU * k)] = fhu[c][k] * jac;
#pragma unroll
      for (int k = 0; k < NCU; ++k) sGHh[t + NFG * (c + NCU * k)] = fhh[c][k] * jac;
    }
  }
  __syncthreads();

  const int32_t* loc = T.loc;   // loc[v + NPV*j]
  const double* sff = T.sff;    // sff[a + NPF*b + NPF*NPF*g]

  // ---- 3a. Ru = shapvg*wrk - sum_faces shapfg*FH (volint.m:258, faceint.m:302-330)
  double* Ru = A.Ru + e * NU;
  for (int idx = tid; idx < NU; idx += NT) {
    const int a = idx % NPV, c = idx / NPV;
    double v = 0.0;
    for (int r = 0; r < NR; ++r) v += T.shapvg[a + NPV * r] * sWrk[r + NR * c];
#pragma unroll
    for (int j = 0; j < NFE; ++j) {
      const int la = loc[a + NPV * j];
      if (la >= 0)
        for (int g = 0; g < NGF; ++g) v -= T.shapfg[la + NPF * g] * sFH[j * NGF + g + NFG * c];
    }
    Ru[idx] = v;
  }
  // ---- 3b. Rh = -shapfg * (fbou | fhat) (faceint.m:335-339)
  double* Rh = A.Rh + e * NH;
  for (int idx = tid; idx < NH; idx += NT) {
    const int c = idx % NCU, ab = idx / NCU, j = ab / NPF, a = ab - j * NPF;
    double v = 0.0;
    for (int g = 0; g < NGF; ++g) v += T.shapfg[a + NPF * g] * sGH[j * NGF + g + NFG * c];
    Rh[idx] = -v;
  }
  // ---- 3c. BD (npv,ncu,npv,nc) = volume + face blocks (volint.m:262, faceint.m:316,326)
  double* BD = A.BD + e * (int64_t)NU * C::NCOL;
  for (int idx = tid; idx < NU * C::NCOL; idx += NT) {
    const int jj = idx % NPV;
    int rest = idx / NPV;
    const int c = rest % NCU;
    rest /= NCU;
    const int ii = rest % NPV;
    const int k = rest / NPV;
    const double* sv = T.svv + jj + NPV * ii;
    const double* wl = sWrl + NR * (c + NCU * k);
    double v = 0.0;
    for (int r = 0; r < NR; ++r) v += sv[NPV * NPV * r] * wl[r];
#pragma unroll
    for (int j = 0; j < NFE; ++j) {
      const int la = loc[jj + NPV * j], lb = loc[ii + NPV * j];
      if (la >= 0 && lb >= 0)
        for (int g = 0; g < NGF; ++g)
          v += sff[la + NPF * lb + NPF * NPF * g] * sFHu[j * NGF + g + NFG * (c + NCU * k)];
    }
    BD[idx] = v;
  }
  // ---- 3d. F (npv,ncu,ncu,ndf) (faceint.m:320,327)
  double* F = A.F + e * (int64_t)NU * NH;
  for (int idx = tid; idx < NU * NH; idx += NT) {
    const int va = idx % NPV;
    int rest = idx / NPV;
    const int c = rest % NCU;
    rest /= NCU;
    const int k = rest % NCU;
    const int bj = rest / NCU, j = bj / NPF, b = bj - j * NPF;
    const int la = loc[va + NPV * j];
    double v = 0.0;
    if (la >= 0)
      for (int g = 0; g < NGF; ++g) v += sff[la + NPF * b + NPF * NPF * g] * sFHh[j * NGF + g + NFG * (c + NCU * k)];
    F[idx] = v;
  }
  // ---- 3e. GK (ncu,ndf,npv,nc) (faceint.m:342-358)
  double* GK = A.GK + e * (int64_t)NH * C::NCOL;
  for (int idx = tid; idx < NH * C::NCOL; idx += NT) {
    const int c = idx % NCU;
    int rest = idx / NCU;
    const int bj = rest % NDF;
    rest /= NDF;
    const int va = rest % NPV, k = rest / NPV;
    const int j = bj / NPF, b = bj - j * NPF;
    const int la = loc[va + NPV * j];
    double v = 0.0;
    if (la >= 0)
      for (int g = 0; g < NGF; ++g) v += sff[la + NPF * b + NPF * NPF * g] * sGHu[j * NGF + g + NFG * (c + NCU * k)];
    GK[idx] = v;
  }
  // ---- 3f. H (ncu,ndf,ncu,ndf), block diagonal per face (faceint.m:347-359)
  double* H = A.H + e * (int64_t)NH * NH;
  for (int idx = tid; idx < NH * NH; idx += NT) {
    const int c = idx % NCU;
    int rest = idx / NCU;
    const int aj = rest % NDF;
    rest /= NDF;
    const int k = rest % NCU, bj = rest / NCU;
    const int j = aj / NPF, a = aj - j * NPF, j2 = bj / NPF, b = bj - j2 * NPF;
    double v = 0.0;
    if (j == j2)
      for (int g = 0; g < NGF; ++g) v += sff[a + NPF * b + NPF * NPF * g] * sGHh[j * NGF + g + NFG * (c + NCU * k)];
    H[idx] = v;
  }
  if constexpr (UQ) {
    // ---- 3g. M, C (volint.m:43-55), E (faceint.m:34-47)
    for (int idx = tid; idx < NPV * NPV; idx += NT) {
      double v = 0.0;
      for (int g = 0; g < NGV; ++g) v += T.svv[idx + NPV * NPV * g] * sJac[g];
      v *= ap.fc_q;
      sM[idx] = v;
      A.M[e * NPV * NPV + idx] = v;
    }
    for (int idx = tid; idx < NPV * NPV * ND; idx += NT) {
      const int jjii = idx % (NPV * NPV), i = idx / (NPV * NPV);
      double v = 0.0;
#pragma unroll
      for (int j = 0; j < ND; ++j)
        for (int g = 0; g < NGV; ++g) v += T.svv[jjii + NPV * NPV * (g + NGV * (j + 1))] * sXx[g * ND * ND + i * ND + j];
      sC[idx] = v;
      A.C[e * NPV * NPV * ND + idx] = v;
    }
    for (int idx = tid; idx < NPV * NDF * ND; idx += NT) {
      const int va = idx % NPV;
      int rest = idx / NPV;
      const int bj = rest % NDF, i = rest / NDF;
      const int j = bj / NPF, b = bj - j * NPF;
      const int la = loc[va + NPV * j];
      double v = 0.0;
      if (la >= 0)
        for (int g = 0; g < NGF; ++g) v += sff[la + NPF * b + NPF * NPF * g] * sNJ[j * NGF + g + NFG * i];
      sE[idx] = v;
      A.E[e * NPV * NDF * ND + idx] = v;
    }
    __syncthreads();
    // ---- 3h. Rq = M (SH_q/fc_q - q) + C u - E uh (volint.m:57-60, faceint.m:48-51)
    for (int idx = tid; idx < NPV * NCU * ND; idx += NT) {
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
}

}  // namespace hdgx
