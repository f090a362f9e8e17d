// Batched Schur elimination of the element unknowns onto the trace unknowns
// (kernel/schur_u.c:42-163, kernel/schur_uq.c:40-287) on gfx950.
//
// One wavefront per element.  Lane i owns row i of D (D + B M^-1 C for u,q) in
// registers; LU with partial pivoting is done by rows: the pivot is a
// wavefront argmax, the pivot row is broadcast through LDS and row interchanges
// are virtual (each lane records the elimination step at which its row became
// the pivot), so no register is ever indexed dynamically.  The right-hand
// sides [Ru, -F] (or [Ru - B M^-1 Rq, B M^-1 E - F]) are solved in register
// chunks of CH columns by forward/back substitution; each solved chunk is
// staged in LDS, extended by the q rows (M^-1 C du, M^-1 C dlu - M^-1 E) and
// multiplied by GK to give the matching columns of AE = H + GK*X and
// FE = Rh - GK*x (schur_uq.c:279-285).
//
// prec = 1 factors and solves in FP32 and refines every chunk against the FP64
// matrix (two refinement sweeps, residual in FP64).
#include <hip/hip_runtime.h>

#include "hdgx_device.h"
#include "hdgx_kernels.h"

namespace hdgx {

enum { ST_NONSPD_M = 1, ST_SINGULAR_D = 2, ST_NONFINITE = 4 };

template <class C, int CH, typename R>
__global__ __launch_bounds__(64) void schur_kernel(SchurArgs A) {
  constexpr int NU = C::NU, NH = C::NH, NCOL = C::NCOL, NPV = C::NPV, NCU = C::NCU, ND = C::ND;
  constexpr int NDF = C::NDF;
  constexpr bool UQ = C::UQ;
  constexpr bool MIXED = sizeof(R) == 4;
  constexpr int NRHS = 1 + NH;
  constexpr int NQ = UQ ? NPV * NCU * ND : 0;
  constexpr int NRQ = NCU * ND, NCC = NPV * ND, NEE = NDF * ND;
  static_assert(NU <= 64, "one row per lane");

  const int lane = threadIdx.x;
  const int64_t e = blockIdx.x;
  if (e >= A.ne) return;

  __shared__ double sPiv[NU];
  __shared__ double sB[CH];
  __shared__ int sPl[NU];
  __shared__ double sX[NCOL * CH];
  __shared__ double sMf[UQ ? NPV * NPV : 1];
  __shared__ double sMiRq[UQ ? NPV * NRQ : 1];
  __shared__ double sMiC[UQ ? NPV * NCC : 1];
  __shared__ double sMiE[UQ ? NPV * NEE : 1];

  int st = 0;
  const double* BDe = A.BD + e * (int64_t)NU * NCOL;

  // ---- u,q prologue: M = U'U (dpotrf 'U'), then M^-1 [Rq C E] (dpotrs)
  if constexpr (UQ) {
    for (int i = lane; i < NPV * NPV; i += 64) sMf[i] = A.M[e * NPV * NPV + i];
    __syncthreads();
    for (int j = 0; j < NPV; ++j) {
      if (lane == 0) {
        double v = sMf[j + NPV * j];
        if (!(v > 0.0)) st |= ST_NONSPD_M;
        sMf[j + NPV * j] = sqrt(v);
      }
      __syncthreads();
      if (lane > j && lane < NPV) sMf[j + NPV * lane] /= sMf[j + NPV * j];
      __syncthreads();
      const int m = NPV - j - 1;
      for (int idx = lane; idx < m * m; idx += 64) {
        const int i = j + 1 + idx % m, k = j + 1 + idx / m;
        if (i <= k) sMf[i + NPV * k] -= sMf[j + NPV * i] * sMf[j + NPV * k];
      }
      __syncthreads();
    }
    for (int col = lane; col < NRQ + NCC + NEE; col += 64) {
      const double* src;
      double* dst;
      if (col < NRQ) { src = A.Rq + e * NPV * NRQ + NPV * col; dst = sMiRq + NPV * col; }
      else if (col < NRQ + NCC) { src = A.C + e * NPV * NCC + NPV * (col - NRQ); dst = sMiC + NPV * (col - NRQ); }
      else { src = A.E + e * NPV * NEE + NPV * (col - NRQ - NCC); dst = sMiE + NPV * (col - NRQ - NCC); }
      double x[NPV];
#pragma unroll
      for (int i = 0; i < NPV; ++i) x[i] = src[i];
#pragma unroll
      for (int i = 0; i < NPV; ++i) {
        double s = x[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= sMf[k + NPV * i] * x[k];
        x[i] = s / sMf[i + NPV * i];
      }
#pragma unroll
      for (int i = NPV - 1; i >= 0; --i) {
        double s = x[i];
#pragma unroll
        for (int k = i + 1; k < NPV; ++k) s -= sMf[i + NPV * k] * x[k];
        x[i] = s / sMf[i + NPV * i];
      }
#pragma unroll
      for (int i = 0; i < NPV; ++i) dst[i] = x[i];
    }
    __syncthreads();
  }

  // ---- the row of D' = D (+ B M^-1 C) and rhs0 = Ru (- B M^-1 Rq) owned by this lane
  const bool row = lane < NU;
  double d64[NU];
  double r0 = 0.0;
  if (row) {
#pragma unroll
    for (int j = 0; j < NU; ++j) d64[j] = BDe[lane + NU * j];
    r0 = A.Ru[e * NU + lane];
    if constexpr (UQ) {
#pragma unroll
      for (int m = 0; m < NCU; ++m)
#pragma unroll
        for (int dd = 0; dd < ND; ++dd) {
          double bv[NPV];
#pragma unroll
          for (int a = 0; a < NPV; ++a) bv[a] = BDe[lane + NU * (NU + a + NPV * m + NPV * NCU * dd)];
#pragma unroll
          for (int b = 0; b < NPV; ++b) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < NPV; ++a) s += bv[a] * sMiC[a + NPV * b + NPV * NPV * dd];
            d64[b + NPV * m] += s;
          }
          double s = 0.0;
#pragma unroll
          for (int a = 0; a < NPV; ++a) s += bv[a] * sMiRq[a + NPV * m + NPV * NCU * dd];
          r0 -= s;
        }
    }
  } else {
#pragma unroll
    for (int j = 0; j < NU; ++j) d64[j] = 0.0;
  }

  // ---- LU with partial pivoting; row i keeps multipliers left of its pivot step
  R d[NU];
#pragma unroll
  for (int j = 0; j < NU; ++j) d[j] = (R)d64[j];
  int mystep = NU;
#pragma unroll
  for (int k = 0; k < NU; ++k) {
    double v = -1.0;
    if (row && mystep == NU) { double a = fabs((double)d[k]); v = isnan(a) ? 1.0e308 : a; }
    int pl = lane;
    wave_argmax(v, pl);
    if (!(v > 0.0)) st |= ST_SINGULAR_D;
    if (lane == pl) {
      mystep = k;
      sPl[k] = lane;
#pragma unroll
      for (int j = k; j < NU; ++j) sPiv[j] = (double)d[j];
    }
    __syncthreads();
    if (row && mystep == NU) {
      const R l = d[k] / (R)sPiv[k];
      d[k] = l;
#pragma unroll
      for (int j = k + 1; j < NU; ++j) d[j] -= l * (R)sPiv[j];
    }
    __syncthreads();
  }
  const double* Fe = A.F + e * (int64_t)NU * NH;
  const double* GKe = A.GK + e * (int64_t)NH * NCOL;
  const double* He = A.H + e * (int64_t)NH * NH;
  double* Xe = A.dudg_duh + e * (int64_t)NCOL * NH;

  for (int c0 = 0; c0 < NRHS; c0 += CH) {
    // right-hand sides of this chunk
    double rhs[CH];
#pragma unroll
    for (int cc = 0; cc < CH; ++cc) {
      const int col = c0 + cc;
      double v = 0.0;
      if (row && col < NRHS) {
        if (col == 0) v = r0;
        else {
          v = -Fe[lane + NU * (col - 1)];
          if constexpr (UQ) {
            // + BMiE(i, jf, k) = sum_dd sum_a B(i, a, k, dd) MiE(a, jf, dd)
            const int k = (col - 1) % NCU, jf = (col - 1) / NCU;
            double s = 0.0;
#pragma unroll
            for (int dd = 0; dd < ND; ++dd)
#pragma unroll
              for (int a = 0; a < NPV; ++a)
                s += BDe[lane + NU * (NU + a + NPV * k + NPV * NCU * dd)] * sMiE[a + NPV * jf + NPV * NDF * dd];
            v += s;
          }
        }
      }
      rhs[cc] = v;
    }

    // solve D' x = rhs with the factors (+ refinement for the mixed form)
    double x[CH];
#pragma unroll
    for (int cc = 0; cc < CH; ++cc) x[cc] = 0.0;
    constexpr int NSWEEP = MIXED ? 3 : 1;
#pragma unroll 1
    for (int sweep = 0; sweep < NSWEEP; ++sweep) {
      R b[CH];
      if (sweep == 0) {
#pragma unroll
        for (int cc = 0; cc < CH; ++cc) b[cc] = (R)rhs[cc];
      } else if constexpr (MIXED) {
        // FP64 residual r = rhs - D' x, x staged in LDS by solution row
        if (row && mystep < NU)
#pragma unroll
          for (int cc = 0; cc < CH; ++cc) sX[mystep + NCOL * cc] = x[cc];
        __syncthreads();
#pragma unroll
        for (int cc = 0; cc < CH; ++cc) {
          double s = rhs[cc];
#pragma unroll
          for (int j = 0; j < NU; ++j) s -= d64[j] * sX[j + NCOL * cc];
          b[cc] = row ? (R)s : (R)0;
        }
        __syncthreads();
      }
      // forward substitution with the unit-lower multipliers
#pragma unroll
      for (int k = 0; k < NU; ++k) {
        const int pl = sPl[k];
        if (lane == pl)
#pragma unroll
          for (int cc = 0; cc < CH; ++cc) sB[cc] = (double)b[cc];
        __syncthreads();
        if (row && mystep > k) {
          const R l = d[k];
#pragma unroll
          for (int cc = 0; cc < CH; ++cc) b[cc] -= l * (R)sB[cc];
        }
        __syncthreads();
      }
      // back substitution with U
#pragma unroll
      for (int k = NU - 1; k >= 0; --k) {
        const int pl = sPl[k];
        if (lane == pl) {
          const R inv = (R)1 / d[k];
#pragma unroll
          for (int cc = 0; cc < CH; ++cc) { b[cc] *= inv; sB[cc] = (double)b[cc]; }
        }
        __syncthreads();
        if (row && mystep < k) {
          const R u = d[k];
#pragma unroll
          for (int cc = 0; cc < CH; ++cc) b[cc] -= u * (R)sB[cc];
        }
        __syncthreads();
      }
#pragma unroll
      for (int cc = 0; cc < CH; ++cc) x[cc] += (double)b[cc];
    }

    // stage the u rows of the solution, then the q rows
    if (row && mystep < NU)
#pragma unroll
      for (int cc = 0; cc < CH; ++cc) sX[mystep + NCOL * cc] = x[cc];
    __syncthreads();
    if constexpr (UQ) {
      for (int rr = lane; rr < NQ; rr += 64) {
        const int n = rr % NPV, m = (rr / NPV) % NCU, dd = rr / (NPV * NCU);
#pragma unroll
        for (int cc = 0; cc < CH; ++cc) {
          const int col = c0 + cc;
          double v = 0.0;
#pragma unroll
          for (int a = 0; a < NPV; ++a) v += sMiC[n + NPV * a + NPV * NPV * dd] * sX[a + NPV * m + NCOL * cc];
          if (col == 0) v += sMiRq[n + NPV * m + NPV * NCU * dd];
          else if (col < NRHS) {
            const int k = (col - 1) % NCU, jf = (col - 1) / NCU;
            if (m == k) v -= sMiE[n + NPV * jf + NPV * NDF * dd];
          }
          sX[NU + rr + NCOL * cc] = v;
        }
      }
      __syncthreads();
    }
    // DUDG / DUDG_DUH rows
    for (int r = lane; r < NCOL; r += 64) {
#pragma unroll
      for (int cc = 0; cc < CH; ++cc) {
        const int col = c0 + cc;
        if (col >= NRHS) break;
        const double v = sX[r + NCOL * cc];
        if (!isfinite(v)) st |= ST_NONFINITE;
        if (col == 0) A.dudg[e * NCOL + r] = v;
        else Xe[r + (int64_t)NCOL * (col - 1)] = v;
      }
    }
    // AE = H + GK X, FE = Rh - GK x
    for (int r = lane; r < NH; r += 64) {
      double acc[CH];
#pragma unroll
      for (int cc = 0; cc < CH; ++cc) acc[cc] = 0.0;
      for (int k = 0; k < NCOL; ++k) {
        const double g = GKe[r + NH * k];
#pragma unroll
        for (int cc = 0; cc < CH; ++cc) acc[cc] += g * sX[k + NCOL * cc];
      }
#pragma unroll
      for (int cc = 0; cc < CH; ++cc) {
        const int col = c0 + cc;
        if (col >= NRHS) break;
        double v;
        if (col == 0) { v = A.Rh[e * NH + r] - acc[cc]; A.fe[e * NH + r] = v; }
        else { v = He[r + NH * (col - 1)] + acc[cc]; A.ae[e * (int64_t)NH * NH + r + NH * (col - 1)] = v; }
        if (!isfinite(v)) st |= ST_NONFINITE;
      }
    }
    __syncthreads();
  }
  st = wave_or(st);
  if (lane == 0 && A.status) A.status[e] = st;
}

}  // namespace hdgx
