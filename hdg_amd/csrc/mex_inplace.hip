// The MEX files' in-place side effects, for callers that compare against them
// or rely on them (HDGX_INPLACE).  kernel/schur_uq.c:localsolveuq leaves in
// its inputs, per element:
//   M  -> dpotrf('U') factor U (M = U'U) in the upper triangle, the strict
//         lower triangle untouched                              (schur_uq.c:69)
//   Rq, C, E -> M^-1 Rq, M^-1 C, M^-1 E (dpotrs)                 (:72-78)
//   Ru -> Ru - B M^-1 Rq                                         (:85-87, 115-116)
//   F  -> B M^-1 E - F, with F's (k, j) blocks in BMiE's order   (:118-125)
//   D  -> dgetrf's LU of D + B M^-1 C, rows interchanged as it pivots
//         (partial pivoting on the first largest |.|)            (:112-113, 136)
// and kernel/schur_u.c:localsolve leaves D = BD -> its LU (schur_u.c:67).
// B is the q part of BD (columns npv*ncu ..).  The elimination kernels never
// write their inputs; this kernel runs after them on the same stream and
// recomputes these values in FP64.  One workgroup per element, sizes at run
// time, everything staged in (dynamic) LDS.
#include <hip/hip_runtime.h>

#include "hdgx_kernels.h"

namespace hdgx {

constexpr int INPLACE_THREADS = 256;

__global__ __launch_bounds__(INPLACE_THREADS) void mex_inplace_kernel(InplaceArgs A) {
  extern __shared__ double lds[];
  const int64_t e = blockIdx.x;
  if (e >= A.ne) return;
  const int tid = threadIdx.x, NT = INPLACE_THREADS;
  const int npv = A.npv, ncu = A.ncu, nd = A.nd, ndf = A.ndf;
  const int na = npv * ncu;                       // rows of D
  const int ncolbd = A.uq ? npv * ncu * (nd + 1) : na;
  double* BD = A.BD + e * (int64_t)na * ncolbd;
  double* sD = lds;                               // D' (na x na), column-major
  __shared__ int piv;

  if (A.uq) {
    const int nrc = npv * nd, nre = ndf * nd, nrq = ncu * nd, nr = nrc + nre + nrq;
    double* sU = sD + na * na;                    // M, then its factor (npv x npv)
    double* sX = sU + npv * npv;                  // [C | E | Rq] (npv x nr), then M^-1 [C | E | Rq]
    double* M = A.M + e * (int64_t)npv * npv;
    double* C = A.C + e * (int64_t)npv * nrc;
    double* E = A.E + e * (int64_t)npv * nre;
    double* Rq = A.Rq + e * (int64_t)npv * nrq;
    double* Ru = A.Ru + e * (int64_t)na;
    double* F = A.F + e * (int64_t)na * ncu * ndf;
    for (int i = tid; i < npv * npv; i += NT) sU[i] = M[i];
    for (int i = tid; i < npv * nrc; i += NT) sX[i] = C[i];
    for (int i = tid; i < npv * nre; i += NT) sX[npv * nrc + i] = E[i];
    for (int i = tid; i < npv * nrq; i += NT) sX[npv * (nrc + nre) + i] = Rq[i];
    __syncthreads();
    // Cholesky, upper, column by column (dpotf2's order)
    for (int j = 0; j < npv; ++j) {
      if (tid == 0) {
        double d = sU[j + npv * j];
        for (int k = 0; k < j; ++k) d -= sU[k + npv * j] * sU[k + npv * j];
        sU[j + npv * j] = sqrt(d);
      }
      __syncthreads();
      for (int i = j + 1 + tid; i < npv; i += NT) {
        double v = sU[j + npv * i];
        for (int k = 0; k < j; ++k) v -= sU[k + npv * j] * sU[k + npv * i];
        sU[j + npv * i] = v / sU[j + npv * j];
      }
      __syncthreads();
    }
    for (int i = tid; i < npv * npv; i += NT)
      if (i % npv <= i / npv) M[i] = sU[i];
    // M^-1 [C | E | Rq]: U' y = b, then U x = y, one column per thread
    for (int r = tid; r < nr; r += NT) {
      double* x = sX + npv * r;
      for (int j = 0; j < npv; ++j) {
        double v = x[j];
        for (int k = 0; k < j; ++k) v -= sU[k + npv * j] * x[k];
        x[j] = v / sU[j + npv * j];
      }
      for (int j = npv - 1; j >= 0; --j) {
        double v = x[j];
        for (int k = j + 1; k < npv; ++k) v -= sU[j + npv * k] * x[k];
        x[j] = v / sU[j + npv * j];
      }
    }
    __syncthreads();
    for (int i = tid; i < npv * nrc; i += NT) C[i] = sX[i];
    for (int i = tid; i < npv * nre; i += NT) E[i] = sX[npv * nrc + i];
    for (int i = tid; i < npv * nrq; i += NT) Rq[i] = sX[npv * (nrc + nre) + i];
    const double* Bq = BD + (int64_t)na * na;     // B: column l + npv*j + npv*ncu*i
    const double* MiC = sX;
    const double* MiE = sX + npv * nrc;
    const double* MiR = sX + npv * (nrc + nre);
    // D' = D + B M^-1 C: column n' + npv*j sums over the q columns of component j
    for (int w = tid; w < na * na; w += NT) {
      const int r = w % na, c = w / na, n2 = c % npv, j = c / npv;
      double v = BD[w];
      for (int i = 0; i < nd; ++i)
        for (int l = 0; l < npv; ++l) v += Bq[r + (int64_t)na * (l + npv * j + na * i)] * MiC[l + npv * (n2 + npv * i)];
      sD[w] = v;
    }
    // Ru - B M^-1 Rq
    for (int r = tid; r < na; r += NT) {
      double v = 0.0;
      for (int col = 0; col < na * nd; ++col) v += Bq[r + (int64_t)na * col] * MiR[col];
      Ru[r] -= v;
    }
    // F(r, k, jf) = sum_i sum_l B(r, l + npv*k + npv*ncu*i) M^-1E(l, jf + ndf*i) - F(r, k, jf)
    for (int w = tid; w < na * ncu * ndf; w += NT) {
      const int r = w % na, rest = w / na, k = rest % ncu, jf = rest / ncu;
      double v = 0.0;
      for (int i = 0; i < nd; ++i)
        for (int l = 0; l < npv; ++l) v += Bq[r + (int64_t)na * (l + npv * k + na * i)] * MiE[l + npv * (jf + ndf * i)];
      F[w] = v - F[w];
    }
  } else {
    for (int i = tid; i < na * na; i += NT) sD[i] = BD[i];
  }
  __syncthreads();

  // LU with partial pivoting, whole rows interchanged (dgetf2's order)
  for (int k = 0; k < na; ++k) {
    if (tid == 0) {
      int p = k;
      double best = fabs(sD[k + na * k]);
      for (int r = k + 1; r < na; ++r) {
        const double a = fabs(sD[r + na * k]);
        if (a > best) { best = a; p = r; }
      }
      piv = p;
    }
    __syncthreads();
    const int p = piv;
    if (p != k)
      for (int c = tid; c < na; c += NT) {
        const double t = sD[k + na * c];
        sD[k + na * c] = sD[p + na * c];
        sD[p + na * c] = t;
      }
    __syncthreads();
    const double pv = sD[k + na * k];
    if (pv != 0.0)
      for (int r = k + 1 + tid; r < na; r += NT) sD[r + na * k] /= pv;
    __syncthreads();
    const int m = na - 1 - k;
    for (int w = tid; w < m * m; w += NT) {
      const int r = k + 1 + w % m, c = k + 1 + w / m;
      sD[r + na * c] -= sD[r + na * k] * sD[k + na * c];
    }
    __syncthreads();
  }
  for (int i = tid; i < na * na; i += NT) BD[i] = sD[i];
}

hipError_t inplace_launch(hipStream_t s, const InplaceArgs& a) {
  if (a.ne <= 0) return hipSuccess;
  const int na = a.npv * a.ncu;
  const int nr = a.npv * a.nd + a.ndf * a.nd + a.ncu * a.nd;
  const size_t words = (size_t)na * na + (a.uq ? (size_t)a.npv * a.npv + (size_t)a.npv * nr : 0);
  hipLaunchKernelGGL(mex_inplace_kernel, dim3((unsigned)a.ne), dim3(INPLACE_THREADS), words * sizeof(double), s, a);
  return hipGetLastError();
}

}  // namespace hdgx
