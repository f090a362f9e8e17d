// q recovery of the HDG local problem on gfx950 (kernel/getq.m:1-129): for
// every element, the q that zeroes the local residual Rq of volint.m:57-60 /
// faceint.m:48-51,
//   q = (fc_q M)^-1 (C u - E uh) + SH_q / fc_q,
// and optionally the nodal gradient qq of getq.m:62-76.
//
// getq.m forms Mi = inv(M)/fc_q, C and E per element and multiplies; here one
// workgroup per element evaluates C u and E uh matrix-free at the Gauss points
// (sum_g shapvg(a,g,j+1) Xx(g,i,j) u_c(g) and sum_g shapfg(a,g) n_i jacf uh_c(g)),
// so neither C (npv^2 nd) nor E (npv ndf nd) is built, and solves the npv x npv
// mass system in LDS by Gaussian elimination (M is symmetric positive
// definite: no pivoting).  getq.m's geometry has the opposite sign convention
// for Xx and E (getq.m:133-143, 99); both flips cancel in q.
//
// Phases (each a flat task list over the workgroup, a barrier between):
//   0. stage the element's nodes, u columns, gathered trace and the small
//      basis tables in LDS;
//   1. volume and face geometry, u at the volume points, uh at the face points;
//   2. [fc_q M | C u - E uh], one task per entry;
//   3. forward elimination, one step (barrier) per pivot, then back
//      substitution likewise;
//   4. q = x + SH_q/fc_q to qdg (element stride ldq, so the q columns of UDG
//      can be the destination), and qq.
//
// Included after quad_kernel.hip (QUAD_THREADS and the geometry helpers).
#include <hip/hip_runtime.h>

#include "hdgx_device.h"
#include "hdgx_kernels.h"

namespace hdgx {

// One wavefront per element: the per-element work is a chain of small
// barrier-separated steps, so many resident elements (not wide ones) hide the
// latency.
constexpr int GETQ_THREADS = 64;

template <class C>
struct GetqLds {
  static constexpr int ND = C::ND, NPV = C::NPV, NPF = C::NPF, NFE = C::NFE, NDF = C::NDF;
  static constexpr int NGV = C::NGV, NGF = C::NGF, NCU = C::NCU, NFG = NFE * NGF, NQ = NCU * ND;
  static constexpr int NA = NPV + NQ;   // columns of [M | rhs]
  double sX[NPV * ND];
  double sU[NPV * NCU];
  double sUH[NCU * NDF];
  double sShV[NGV * NPV];
  double sShF[NGF * NPF];
  double sJac[NGV];
  double sXx[NGV * ND * ND];
  double sNJ[NFG * ND];                 // n_i * jacf at the face points
  double sUg[NGV * NCU];
  double sUHf[NFG * NCU];
  double sA[NPV * NA];                  // [fc_q M | rhs], column-major, leading dimension NPV
  double sJi[NPV * ND * ND];            // -J^-1 at the nodes (qq)
  int32_t sLoc[NPV * NFE];
};

template <class C>
__global__ __launch_bounds__(GETQ_THREADS) void getq_kernel(GetqArgs A, DevMaster T) {
  using L = GetqLds<C>;
  constexpr int ND = L::ND, NPV = L::NPV, NPF = L::NPF, NFE = L::NFE, NDF = L::NDF;
  constexpr int NGV = L::NGV, NGF = L::NGF, NCU = L::NCU, NFG = L::NFG, NQ = L::NQ, NA = L::NA;
  constexpr int NH = NCU * NDF, NT = GETQ_THREADS;
  __shared__ L lds;
  double* const sX = lds.sX;
  double* const sU = lds.sU;
  double* const sUH = lds.sUH;
  double* const sShV = lds.sShV;
  double* const sShF = lds.sShF;
  double* const sJac = lds.sJac;
  double* const sXx = lds.sXx;
  double* const sNJ = lds.sNJ;
  double* const sUg = lds.sUg;
  double* const sUHf = lds.sUHf;
  double* const sA = lds.sA;
  double* const sJi = lds.sJi;
  int32_t* const sLoc = lds.sLoc;

  const int64_t e = blockIdx.x;
  if (e >= A.ne) return;
  const int tid = threadIdx.x;
  const int nc = A.nc, ncd = A.ncd;

  // ---- 0. stage
  for (int i = tid; i < NPV * ND; i += NT) sX[i] = A.dgnodes[e * NPV * ncd + i];
  for (int i = tid; i < NPV * NCU; i += NT) sU[i] = A.udg[e * NPV * nc + i];
  for (int i = tid; i < NH; i += NT) sUH[i] = A.uh[e * NH + i];
  for (int i = tid; i < NGV * NPV; i += NT) sShV[i] = T.shapvt[i];
  for (int i = tid; i < NGF * NPF; i += NT) sShF[i] = T.shapft[i];
  for (int i = tid; i < NPV * NFE; i += NT) sLoc[i] = T.loc[i];
  __syncthreads();

  // ---- 1. geometry and the state at the quadrature points
  {
    constexpr int W0 = NGV, W1 = W0 + NFG, W2 = W1 + NGV * NCU, W3 = W2 + NFG * NCU;
    const int W4 = W3 + (A.qq ? NPV : 0);
    for (int w = tid; w < W4; w += NT) {
      if (w < W0) {
        double Xx[ND][ND];
        sJac[w] = vol_geom<C>(T, sX, w, Xx);
#pragma unroll
        for (int a = 0; a < ND; ++a)
#pragma unroll
          for (int b = 0; b < ND; ++b) sXx[w * ND * ND + a * ND + b] = Xx[a][b];
      } else if (w < W1) {
        const int t = w - W0, j = t / NGF, g = t - j * NGF;
        double nl[ND];
        const double jac = face_geom<C>(T, sX, j, g, nl);
#pragma unroll
        for (int i = 0; i < ND; ++i) sNJ[t + NFG * i] = nl[i] * jac;
      } else if (w < W2) {
        const int i = w - W1, g = i % NGV, c = i / NGV;
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < NPV; ++a) v += sShV[g + NGV * a] * sU[a + NPV * c];
        sUg[i] = v;
      } else if (w < W3) {
        const int i = w - W2, t = i % NFG, c = i / NFG;
        const int j = t / NGF, g = t - j * NGF;
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < NPF; ++a) v += sShF[g + NGF * a] * sUH[c + NCU * (a + NPF * j)];
        sUHf[i] = v;
      } else {
        // node k: -J^-1 from the nodal basis derivatives (getq.m:62-70)
        const int k = w - W3;
        double J[ND][ND], Xx[ND][ND];
#pragma unroll
        for (int a = 0; a < ND; ++a)
#pragma unroll
          for (int b = 0; b < ND; ++b) {
            double v = 0.0;
            for (int f = 0; f < NPV; ++f) v += A.shapnv[f + NPV * k + NPV * NPV * (1 + a)] * sX[f + NPV * b];
            J[a][b] = v;   // dx_b / dxi_a
          }
        const double ij = -1.0 / adjugate<ND>(J, Xx);
#pragma unroll
        for (int a = 0; a < ND; ++a)
#pragma unroll
          for (int b = 0; b < ND; ++b) sJi[k * ND * ND + a * ND + b] = Xx[a][b] * ij;
      }
    }
  }
  __syncthreads();

  // ---- 2. [fc_q M | C u - E uh]
  for (int w = tid; w < NPV * NA; w += NT) {
    const int a = w % NPV, col = w / NPV;
    double v = 0.0;
    if (col < NPV) {
      // M(a, col) = sum_g shapvg(a,g) shapvl(col,g) jac(g) (volint.m:43)
#pragma unroll
      for (int g = 0; g < NGV; ++g) v += T.svv[w + NPV * NPV * g] * sJac[g];
      v *= A.fc_q;
    } else {
      const int q = col - NPV, c = q % NCU, i = q / NCU;
      // (C u)(a, c, i) = sum_g sum_j shapvg(a,g,j+1) Xx(g,i,j) u_c(g) (volint.m:46-55)
#pragma unroll
      for (int g = 0; g < NGV; ++g) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < ND; ++j) s += T.shapvg[a + NPV * g + NPV * NGV * (1 + j)] * sXx[g * ND * ND + i * ND + j];
        v += s * sUg[g + NGV * c];
      }
      // - (E uh)(a, c, i) over the faces that hold node a (faceint.m:34-51)
#pragma unroll
      for (int j = 0; j < NFE; ++j) {
        const int la = sLoc[a + NPV * j];
        if (la < 0) continue;
        double s = 0.0;
#pragma unroll
        for (int g = 0; g < NGF; ++g) {
          const int t = j * NGF + g;
          s += T.shapfg[la + NPF * g] * sNJ[t + NFG * i] * sUHf[t + NFG * c];
        }
        v -= s;
      }
    }
    sA[w] = v;
  }
  __syncthreads();

  // ---- 3. Gaussian elimination of [M | rhs], then back substitution
  for (int k = 0; k < NPV - 1; ++k) {
    const int nr = NPV - 1 - k, ncl = NA - 1 - k;
    const double pinv = 1.0 / sA[k + NPV * k];
    for (int w = tid; w < nr * ncl; w += NT) {
      const int i = k + 1 + w % nr, col = k + 1 + w / nr;
      sA[i + NPV * col] -= sA[i + NPV * k] * pinv * sA[k + NPV * col];
    }
    __syncthreads();
  }
  for (int k = NPV - 1; k > 0; --k) {
    const double pinv = 1.0 / sA[k + NPV * k];
    for (int w = tid; w < k * NQ; w += NT) {
      const int i = w % k, col = NPV + w / k;
      sA[i + NPV * col] -= sA[i + NPV * k] * pinv * sA[k + NPV * col];
    }
    __syncthreads();
  }

  // ---- 4. outputs
  const double ifq = 1.0 / A.fc_q;
  double* const qe = A.qdg + e * A.ldq;
  for (int w = tid; w < NPV * NQ; w += NT) {
    const int a = w % NPV, q = w / NPV;
    double v = sA[a + NPV * (NPV + q)] / sA[a + NPV * a];
    // SH(:, end-ncq+1:end) / fc_q (getq.m:125-127)
    if (A.sh) v += A.sh[e * NPV * nc + NPV * (nc - NQ) + w] * ifq;
    qe[w] = v;
  }
  if (A.qq) {
    // qq(k, c + ncu*r) = sum_a (-J^-1)(r, a) du_c/dxi_a at node k (getq.m:71-75)
    for (int w = tid; w < NPV * NQ; w += NT) {
      const int k = w % NPV, q = w / NPV, c = q % NCU, r = q / NCU;
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < ND; ++a) {
        double du = 0.0;
        for (int f = 0; f < NPV; ++f) du += A.shapnv[f + NPV * k + NPV * NPV * (1 + a)] * sU[f + NPV * c];
        v += sJi[k * ND * ND + r * ND + a] * du;
      }
      A.qq[e * NPV * NQ + w] = v;
    }
  }
}

}  // namespace hdgx
