// u-only FP64 elimination for the large local systems with four 16-row tiles
// and n_u <= 56 (Euler p = 2 tetrahedra, n_u = 50),
// kernel/schur_u.c:42-80 and its mexFunction's AE/FE (:143-161):
//   [du | dlu] = D \ [Ru | -F],  AE = H + GK dlu,  FE = Rh - GK du.
// The MFMA form holds [D | RHS] in accumulator tiles, which at four row tiles
// needs 256 VGPRs plus 256 AGPRs: one wavefront per SIMD, latency exposed.
// Here one 128-thread workgroup (two wavefronts) per element keeps D in LDS
// and four elements share a CU (243 VGPRs, two wavefronts per SIMD; its phases
// are latency chains, so elements in flight set the rate -- the 256-thread
// form held three at 167 VGPRs and was 4 % slower):
//   1. D -> LDS; LU with partial pivoting (dgetf2's order: the first largest
//      |.| of the column, whole rows interchanged), three barriers per column;
//   2. thread t < 1 + n_h takes RHS column t (Ru, then -F's columns) into
//      registers in pivot order, solves L y = P b and U x = y reading the
//      factors from LDS (every lane the same address: broadcast), and writes
//      it to DUDG (t = 0) or DUDG_DUH;
//   3. GK passes through LDS 8 rows at a time (coalesced loads by the whole
//      workgroup), and thread t forms those 8 entries of FE (t = 0) or of its
//      AE column, 8 accumulators fed by wide broadcast reads.
#include <hip/hip_runtime.h>

#include "hdgx_device.h"
#include "hdgx_kernels.h"

namespace hdgx {

constexpr int SCHUR_LDS_THREADS = 128;   // two wavefronts: the 121 RHS columns of tetrahedra p2

template <class C>
struct SchurLdsShape {
  static constexpr int GR = 8;                        // GK rows per LDS pass (16 spills)
  static constexpr int NU = C::NU, NH = C::NH, LDA = NU | 1;
  static constexpr int NRHS = 1 + NH;
  // n_u <= 56: the solution column (2 n_u VGPRs) fits; at n_u = 60 (Euler
  // p = 4 triangles) the solve spills at one wavefront per SIMD and those
  // systems stay on the MFMA form
  static constexpr bool USE = !C::UQ && NU > 48 && NU <= 56 && NRHS <= SCHUR_LDS_THREADS;
};

// the larger |value|, the smaller row on ties (idamax's choice)
HDGX_DEV void argmax_step(double& v, int& i, double v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

// Element e of the batch A, one SCHUR_LDS_THREADS-thread workgroup.
template <class C>
HDGX_DEV void schur_lds_body(const SchurArgs& A, const int64_t e) {
  using S = SchurLdsShape<C>;
  constexpr int NU = S::NU, NH = S::NH, LDA = S::LDA, NT = SCHUR_LDS_THREADS, GR = S::GR;
  __shared__ double sA[NU * LDA];
  __shared__ double sG[GR * NU];                       // GK rows r0..r0+GR-1, row-fastest
  __shared__ int sPerm[NU];
  __shared__ int sPiv;
  __shared__ int sSt;
  if (e >= A.ne) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

  const double* D = A.BD + e * (int64_t)NU * NU;
  for (int i = tid; i < NU * NU; i += NT) sA[i % NU + LDA * (i / NU)] = D[i];
  for (int i = tid; i < NU; i += NT) sPerm[i] = i;
  if (tid == 0) sSt = 0;
  __syncthreads();

  // ---- 1. LU with partial pivoting in LDS
  for (int k = 0; k < NU; ++k) {
    if (wv == 0) {
      double v = -1.0;
      int i = NU;
      const int r = k + lane;
      if (r < NU) { v = fabs(sA[r + LDA * k]); i = r; }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double v2 = __shfl_xor(v, off, 64);
        const int i2 = __shfl_xor(i, off, 64);
        argmax_step(v, i, v2, i2);
      }
      if (lane == 0) {
        sPiv = i;
        if (v == 0.0) sSt |= 2;
      }
    }
    __syncthreads();
    const int p = sPiv;
    if (p != k) {
      for (int c = tid; c < NU; c += NT) {
        const double t = sA[k + LDA * c];
        sA[k + LDA * c] = sA[p + LDA * c];
        sA[p + LDA * c] = t;
      }
      if (tid == 0) {
        const int t = sPerm[k];
        sPerm[k] = sPerm[p];
        sPerm[p] = t;
      }
    }
    __syncthreads();
    const double pv = sA[k + LDA * k];
    const double ipv = pv != 0.0 ? 1.0 / pv : 0.0;
    const int m = NU - 1 - k;
    // trailing update with the unscaled column (scaled just after)
    for (int w = tid; w < m * m; w += NT) {
      const int i = k + 1 + w % m, c = k + 1 + w / m;
      sA[i + LDA * c] -= sA[i + LDA * k] * ipv * sA[k + LDA * c];
    }
    __syncthreads();
    // multipliers of column k: done before the next step's first barrier,
    // which the row interchange waits for (the pivot search in between reads
    // column k + 1 only)
    if (pv != 0.0)
      for (int i = k + 1 + tid; i < NU; i += NT) sA[i + LDA * k] *= ipv;
  }
  __syncthreads();

  // ---- 2. one RHS column per thread: solve and write
  const int t = tid;
  const bool act = t < S::NRHS;
  int st = 0;
  double x[NU];
  if (act) {
    const double* src = t == 0 ? A.Ru + e * (int64_t)NU : A.F + e * (int64_t)NU * NH + (int64_t)NU * (t - 1);
    const double sg = t == 0 ? 1.0 : -1.0;
#pragma unroll
    for (int i = 0; i < NU; ++i) x[i] = sg * src[sPerm[i]];
    // L y = P b (unit lower)
#pragma unroll
    for (int k = 0; k < NU; ++k)
#pragma unroll
      for (int i = k + 1; i < NU; ++i) x[i] -= sA[i + LDA * k] * x[k];
    // U x = y
#pragma unroll
    for (int k = NU - 1; k >= 0; --k) {
      x[k] /= sA[k + LDA * k];
#pragma unroll
      for (int i = 0; i < k; ++i) x[i] -= sA[i + LDA * k] * x[k];
    }
    const int ldo = C::NPV * A.nco;             // rows of DUDG / DUDG_DUH kept
    if (ldo > 0) {
      double* dst = t == 0 ? A.dudg + e * (int64_t)ldo : A.dudg_duh + e * (int64_t)ldo * NH + (int64_t)ldo * (t - 1);
#pragma unroll
      for (int i = 0; i < NU; ++i)
        if (i < ldo) dst[i] = x[i];
    }
#pragma unroll
    for (int i = 0; i < NU; ++i)
      if (!isfinite(x[i])) st |= 4;
  }

  // ---- 3. AE(:, t-1) = H(:, t-1) + GK x;  FE = Rh - GK x (t = 0)
  const double* GK = A.GK + e * (int64_t)NH * NU;
  const double* in = t == 0 ? A.Rh + e * (int64_t)NH : A.H + e * (int64_t)NH * NH + (int64_t)NH * (t - 1);
  double* out = t == 0 ? A.fe + e * (int64_t)NH : A.ae + e * (int64_t)NH * NH + (int64_t)NH * (t - 1);
  const double sgn = t == 0 ? -1.0 : 1.0;
#pragma unroll 1
  for (int r0 = 0; r0 < NH; r0 += GR) {
    const int nr = NH - r0 < GR ? NH - r0 : GR;
    __syncthreads();                           // the previous rows are consumed
    for (int w = tid; w < GR * NU; w += NT) {
      const int rr = w % GR, k = w / GR;
      sG[w] = rr < nr ? GK[r0 + rr + (int64_t)NH * k] : 0.0;
    }
    __syncthreads();
    if (act) {
      double acc[GR];
#pragma unroll
      for (int rr = 0; rr < GR; ++rr) acc[rr] = 0.0;
#pragma unroll
      for (int k = 0; k < NU; ++k)
#pragma unroll
        for (int rr = 0; rr < GR; ++rr) acc[rr] += sG[rr + GR * k] * x[k];
#pragma unroll
      for (int rr = 0; rr < GR; ++rr)
        if (rr < nr) {
          const double v = in[r0 + rr] + sgn * acc[rr];
          if (!isfinite(v)) st |= 4;
          out[r0 + rr] = v;
        }
    }
  }
  if (st) atomicOr(&sSt, st);
  __syncthreads();
  if (tid == 0 && A.status) A.status[e] = sSt;
}

template <class C>
__global__ __launch_bounds__(SCHUR_LDS_THREADS) void schur_lds_kernel(SchurArgs A) {
  schur_lds_body<C>(A, blockIdx.x);
}

}  // namespace hdgx
