// Trace system of the HDG Newton step on gfx950 (kernel/hdg_solve.m:105-176
// with app.denseblock = 1): face-block assembly, the block operator, a
// multicolour block ILU(0), GMRES(restart) and the local Newton update.
//
// Layout in HBM (all FP64 column-major):
//   UH / x / F     (ncf, nf)                    ncf = ncu*npf, face-major
//   K              (ncf, ncf, nbf-1, nf)        off-diagonal blocks of row i
//   LU             (ncf, ncf, nbf, nf)          slot 0 = inv(U_ii), slot k = L_ik | U_ik
//   Krylov basis V (ncf*nf, restart+1)          in the context's trace workspace
// One workgroup per face for the dense per-face work (assembly, ILU rows);
// one thread per (row, face) for the block operator and the ILU sweeps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <vector>

#include "../../include/hdgx.h"
#include "hdgx_ctx.h"

using hdgx_detail::fail;

namespace {

constexpr int RED_BLOCKS = 512;   // blocks of the first reduction stage
constexpr int RED_THREADS = 256;
constexpr int MAXM = 32;          // dot products per reduction launch (restart + 1 <= MAXM)

__device__ __forceinline__ void wave_argmax_abs(double& v, int& idx) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
}

// Gauss-Jordan elimination with partial pivoting of the leading (n x n) block
// of the column-major (n x ncol) matrix S in LDS, applied to all ncol columns:
// on return columns n.. hold S(:, 0:n)^-1 S(:, n..).  One 64-lane workgroup,
// n <= 64.  fcol: n doubles of LDS scratch.  Returns false on a zero pivot.
__device__ bool lds_gauss_jordan(double* S, int n, int ncol, double* fcol) {
  const int lane = threadIdx.x;
  bool ok = true;
  for (int p = 0; p < n; ++p) {
    double v = (lane >= p && lane < n) ? fabs(S[lane + n * p]) : -1.0;
    int idx = lane;
    wave_argmax_abs(v, idx);
    const int piv = __builtin_amdgcn_readfirstlane(idx);
    if (!(v > 0.0)) ok = false;
    if (piv != p)
      for (int col = lane; col < ncol; col += 64) {
        const double t = S[p + n * col];
        S[p + n * col] = S[piv + n * col];
        S[piv + n * col] = t;
      }
    __syncthreads();
    if (lane < n) fcol[lane] = S[lane + n * p];
    __syncthreads();
    const double ip = 1.0 / fcol[p];
    for (int col = lane; col < ncol; col += 64) S[p + n * col] *= ip;
    __syncthreads();
    for (int idx2 = lane; idx2 < n * ncol; idx2 += 64) {
      const int r = idx2 % n, col = idx2 / n;
      if (r != p && col >= p) S[r + n * col] -= fcol[r] * S[p + n * col];
    }
    __syncthreads();
  }
  return ok;
}

// slot k of face row i: the element it comes from, that element's side
// (0 first, 1 second), and the local face of the block column
__device__ __forceinline__ void slot_info(int k, int nfe, int e1, int j1, int e2, int j2, int& e, int& side, int& lj) {
  if (k == 0) { e = e1; side = 0; lj = j1; return; }
  if (k < nfe) { e = e1; side = 0; lj = (k - 1 < j1) ? k - 1 : k; return; }
  const int kk = k - (nfe - 1);
  e = e2; side = 1; lj = (kk - 1 < j2) ? kk - 1 : kk;
}

// ---- face-block assembly, block-Jacobi scaled (mkDenseBlockSystem.m:35-95)
// Block rows are ordered (c, m): component fastest, then node m within the
// face in the face's own numbering; the element's local face node a maps to
// m = elcon(a, j, e) - i*npf (the element sees the face possibly reversed).
__global__ __launch_bounds__(64) void trace_assemble_kernel(int ncu, int npf, int nfe, int64_t nf,
                                                            const int32_t* __restrict__ fconn,
                                                            const int32_t* __restrict__ f2f,
                                                            const int32_t* __restrict__ elcon,
                                                            const double* __restrict__ AE,
                                                            const double* __restrict__ FE, double* __restrict__ K,
                                                            double* __restrict__ F, int32_t* __restrict__ status) {
  extern __shared__ double lds[];
  const int64_t i = blockIdx.x;
  if (i >= nf) return;
  const int lane = threadIdx.x;
  const int ncf = ncu * npf, nbf = 2 * nfe - 1, nb1 = nbf - 1;
  const int ndf = npf * nfe, nh = ncu * ndf;
  const int ncol = ncf * nbf + 1;                 // [M | K_1 .. K_nb1 | F]
  double* S = lds;
  double* fcol = S + (size_t)ncf * ncol;
  int* imap = reinterpret_cast<int*>(fcol + ncf); // imap[side*npf + m]: element-local face node
  int* cmap = imap + 2 * npf;                     // cmap[k*npf + m'] for the slot's column face
  int* bad = cmap + nbf * npf;

  const int e1 = fconn[4 * i], j1 = fconn[4 * i + 1], e2 = fconn[4 * i + 2], j2 = fconn[4 * i + 3];
  if (lane == 0) *bad = 0;
  for (int t = lane; t < 2 * npf + nbf * npf; t += 64) (t < 2 * npf ? imap[t] : cmap[t - 2 * npf]) = 0;
  __syncthreads();
  if (lane < npf) {
    const int a = lane;
    for (int s = 0; s < 2; ++s) {
      const int es = s ? e2 : e1, js = s ? j2 : j1;
      if (es < 0) continue;
      const int m = elcon[a + npf * js + (int64_t)ndf * es] - (int)(i * npf);
      if (m >= 0 && m < npf) imap[s * npf + m] = a;
      else atomicOr(bad, 1);
    }
    for (int k = 1; k < nbf; ++k) {
      const int jf = f2f[k + (int64_t)nbf * i];
      if (jf < 0) continue;
      int e, side, lj;
      slot_info(k, nfe, e1, j1, e2, j2, e, side, lj);
      const int m = elcon[a + npf * lj + (int64_t)ndf * e] - jf * npf;
      if (m >= 0 && m < npf) cmap[k * npf + m] = a;
      else atomicOr(bad, 1);
    }
  }
  __syncthreads();

  for (int idx = lane; idx < ncf * ncol; idx += 64) {
    const int r = idx % ncf, q = idx / ncf;
    const int c = r % ncu, m = r / ncu;
    double v = 0.0;
    if (q < ncf) {
      const int c2 = q % ncu, m2 = q / ncu;
      for (int s = 0; s < 2; ++s) {
        const int es = s ? e2 : e1, js = s ? j2 : j1;
        if (es < 0) continue;
        const int a = imap[s * npf + m], b = imap[s * npf + m2];
        v += AE[(c + ncu * (a + npf * js)) + (int64_t)nh * (c2 + ncu * (b + npf * js)) + (int64_t)nh * nh * es];
      }
    } else if (q < ncf * nbf) {
      const int k = q / ncf, qq = q % ncf, c2 = qq % ncu, m2 = qq / ncu;
      const int jf = f2f[k + (int64_t)nbf * i];
      if (jf >= 0) {
        int e, side, lj;
        slot_info(k, nfe, e1, j1, e2, j2, e, side, lj);
        const int js = side ? j2 : j1;
        const int a = imap[side * npf + m], b = cmap[k * npf + m2];
        v = AE[(c + ncu * (a + npf * js)) + (int64_t)nh * (c2 + ncu * (b + npf * lj)) + (int64_t)nh * nh * e];
      }
    } else {
      for (int s = 0; s < 2; ++s) {
        const int es = s ? e2 : e1, js = s ? j2 : j1;
        if (es < 0) continue;
        v += FE[c + ncu * (imap[s * npf + m] + npf * js) + (int64_t)nh * es];
      }
    }
    S[idx] = v;
  }
  __syncthreads();
  const bool ok = lds_gauss_jordan(S, ncf, ncol, fcol);
  double* Ki = K + (size_t)ncf * ncf * nb1 * i;
  for (int idx = lane; idx < ncf * ncf * nb1; idx += 64) Ki[idx] = S[idx + ncf * ncf];
  for (int r = lane; r < ncf; r += 64) F[r + (int64_t)ncf * i] = S[r + (size_t)ncf * (ncol - 1)];
  if (lane == 0 && status) status[i] = (ok ? 0 : HDGX_ST_SINGULAR_D) | (*bad ? HDGX_ST_NONFINITE : 0);
}

// ---- y = x + sum_k K_k x(f2f(k)) (gmres.m:38-42, arraymul.c:53-70); NCF/NBF
// fix the block shape at compile time (0: the runtime values)
template <int NCF, int NBF>
// with nrm2: the operand is x / sqrt(*nrm2) (a Krylov vector not yet
// normalised), y = A x / sqrt(*nrm2), and the normalised vector is written to
// vout for this face's own rows
__global__ __launch_bounds__(256) void trace_matvec_kernel(int ncf_, int nbf_, int64_t nf, const double* __restrict__ K,
                                                           const int32_t* __restrict__ f2f,
                                                           const double* __restrict__ x, double* __restrict__ y,
                                                           const double* __restrict__ nrm2, double* __restrict__ vout,
                                                           const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int ncf = NCF ? NCF : ncf_, nbf = NBF ? NBF : nbf_;
  const int fpb = blockDim.x / ncf;
  const int r = threadIdx.x % ncf, fl = threadIdx.x / ncf;
  if (fl >= fpb) return;
  const int64_t i = (int64_t)blockIdx.x * fpb + fl;
  if (i >= nf) return;
  const int nb1 = nbf - 1;
  const double* Ki = K + (size_t)ncf * ncf * nb1 * i;
  double acc = 0.0;
#pragma unroll
  for (int k = 1; k < nbf; ++k) {
    const int j = f2f[k + (int64_t)nbf * i];
    if (j < 0) continue;
    const double* Kk = Ki + (size_t)ncf * ncf * (k - 1);
    const double* xj = x + (size_t)ncf * j;
#pragma unroll
    for (int c = 0; c < ncf; ++c) acc += Kk[r + ncf * c] * xj[c];
  }
  const double xi = x[r + (int64_t)ncf * i];
  if (nrm2) {
    const double sc = 1.0 / sqrt(*nrm2);
    y[r + (int64_t)ncf * i] = (xi + acc) * sc;
    vout[r + (int64_t)ncf * i] = xi * sc;
  } else {
    y[r + (int64_t)ncf * i] = xi + acc;
  }
}

// ---- block ILU(0), one row of the current colour per workgroup
// (computeBILU0.m:6-44 in row-oriented form): for the earlier neighbours k of
// row i in elimination order, L_ik = A_ik U_kk^-1 and A_ij -= L_ik U_kj for
// every later column j of row i that row k also holds; then U_ii^-1.
__global__ __launch_bounds__(64) void bilu0_row_kernel(int ncf, int nbf, const int32_t* __restrict__ rows, int nrows,
                                                       const int32_t* __restrict__ f2f,
                                                       const int32_t* __restrict__ rank, const double* __restrict__ K,
                                                       double* __restrict__ LU, int32_t* __restrict__ status) {
  extern __shared__ double lds[];
  if ((int)blockIdx.x >= nrows) return;
  const int i = rows[blockIdx.x];
  const int lane = threadIdx.x;
  const int nb1 = nbf - 1, n2 = ncf * ncf;
  double* W = lds;                       // (ncf, ncf, nbf) working row
  double* T = W + (size_t)n2 * nbf;      // (ncf, 2*ncf) product / inversion scratch
  double* fcol = T + 2 * n2;
  int* order = reinterpret_cast<int*>(fcol + ncf);   // earlier slots in elimination order
  int* nlow = order + nbf;

  const int ri = rank[i];
  for (int idx = lane; idx < n2 * nbf; idx += 64) {
    const int s = idx / n2, rc = idx % n2;
    double v;
    if (s == 0) v = (rc % ncf == rc / ncf) ? 1.0 : 0.0;
    else v = (f2f[s + (int64_t)nbf * i] >= 0) ? K[rc + (size_t)n2 * (s - 1) + (size_t)n2 * nb1 * i] : 0.0;
    W[idx] = v;
  }
  if (lane == 0) {
    int n = 0;
    for (int s = 1; s < nbf; ++s) {
      const int j = f2f[s + (int64_t)nbf * i];
      if (j >= 0 && rank[j] < ri) {
        int pos = n++;
        while (pos > 0 && rank[f2f[order[pos - 1] + (int64_t)nbf * i]] > rank[j]) { order[pos] = order[pos - 1]; --pos; }
        order[pos] = s;
      }
    }
    *nlow = n;
  }
  __syncthreads();
  const int nl = *nlow;
  for (int o = 0; o < nl; ++o) {
    const int s = order[o];
    const int jk = f2f[s + (int64_t)nbf * i];
    const int rk = rank[jk];
    const double* LUk = LU + (size_t)n2 * nbf * jk;
    // L = W_s U_kk^-1
    for (int idx = lane; idx < n2; idx += 64) {
      const int r = idx % ncf, c = idx / ncf;
      double v = 0.0;
      for (int q = 0; q < ncf; ++q) v += W[r + ncf * q + n2 * s] * LUk[q + ncf * c];
      T[idx] = v;
    }
    __syncthreads();
    for (int idx = lane; idx < n2; idx += 64) W[idx + n2 * s] = T[idx];
    __syncthreads();
    // W_t -= L U_k,j for the later columns j of row i present in row k
    for (int t = 0; t < nbf; ++t) {
      const int jt = (t == 0) ? i : f2f[t + (int64_t)nbf * i];
      if (jt < 0 || rank[jt] <= rk) continue;
      int u = -1;
      for (int q = 1; q < nbf; ++q)
        if (f2f[q + (int64_t)nbf * jk] == jt) u = q;
      if (u < 0) continue;   // fill outside the pattern is dropped
      const double* Ukj = LUk + (size_t)n2 * u;
      for (int idx = lane; idx < n2; idx += 64) {
        const int r = idx % ncf, c = idx / ncf;
        double v = 0.0;
        for (int q = 0; q < ncf; ++q) v += W[r + ncf * q + n2 * s] * Ukj[q + ncf * c];
        W[idx + n2 * t] -= v;
      }
      __syncthreads();
    }
  }
  // U_ii^-1 by Gauss-Jordan on [U_ii | I]
  for (int idx = lane; idx < 2 * n2; idx += 64) {
    const int r = idx % ncf, c = idx / ncf;
    T[idx] = (c < ncf) ? W[idx] : ((c - ncf == r) ? 1.0 : 0.0);
  }
  __syncthreads();
  const bool ok = lds_gauss_jordan(T, ncf, 2 * ncf, fcol);
  double* LUi = LU + (size_t)n2 * nbf * i;
  for (int idx = lane; idx < n2; idx += 64) LUi[idx] = T[idx + n2];
  for (int idx = lane; idx < n2 * nb1; idx += 64) LUi[idx + n2] = W[idx + n2];
  if (lane == 0 && status) status[i] = ok ? 0 : HDGX_ST_SINGULAR_D;
}

// forward sweep over one colour: y_i = x_i - sum_{earlier k} L_ik y_k (applyBILU0.m:12-29)
template <int NCF, int NBF>
__global__ __launch_bounds__(256) void bilu0_lower_kernel(int ncf_, int nbf_, const int32_t* __restrict__ rows, int nrows,
                                                          const int32_t* __restrict__ f2f,
                                                          const int32_t* __restrict__ rank,
                                                          const double* __restrict__ LU, const double* __restrict__ x,
                                                          double* __restrict__ y, int n0, double* __restrict__ vd,
                                                          const int* __restrict__ stop) {
  if (stop && *stop) return;
  __shared__ double tmp[256];
  const int ncf = NCF ? NCF : ncf_, nbf = NBF ? NBF : nbf_;
  const int fpb = blockDim.x / ncf;
  const int r = threadIdx.x % ncf, fl = threadIdx.x / ncf;
  const int64_t ri = (int64_t)blockIdx.x * fpb + fl;
  const bool active = fl < fpb && ri < nrows;
  const int i = active ? rows[ri] : 0;
  const int n2 = ncf * ncf;
  const double* LUi = LU + (size_t)n2 * nbf * i;
  double v = 0.0;
  if (active) {
    const int rk = rank[i];
    v = x[r + (int64_t)ncf * i];
#pragma unroll
    for (int k = 1; k < nbf; ++k) {
      const int j = f2f[k + (int64_t)nbf * i];
      if (j < 0 || rank[j] > rk) continue;
      const double* L = LUi + (size_t)n2 * k;
      // the first colour has no earlier neighbours, so its y is x
      const double* yj = (rank[j] < n0 ? x : y) + (size_t)ncf * j;
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < ncf; ++c) s += L[r + ncf * c] * yj[c];
      v -= s;
    }
  }
  if (!vd) {
    if (active) y[r + (int64_t)ncf * i] = v;
    return;
  }
  // last colour: no later neighbours, so the upper solve is v = U_ii^-1 y
  tmp[threadIdx.x] = v;
  __syncthreads();
  if (active) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < ncf; ++c) s += LUi[r + ncf * c] * tmp[fl * ncf + c];
    vd[r + (int64_t)ncf * i] = s;
  }
}

// backward sweep over one colour: v_i = U_ii^-1 (y_i - sum_{later k} U_ik v_k)
// (applyBILU0.m:31-55); the rows of a face sit in one workgroup
template <int NCF, int NBF>
__global__ __launch_bounds__(256) void bilu0_upper_kernel(int ncf_, int nbf_, const int32_t* __restrict__ rows, int nrows,
                                                          const int32_t* __restrict__ f2f,
                                                          const int32_t* __restrict__ rank,
                                                          const double* __restrict__ LU, const double* __restrict__ y,
                                                          double* __restrict__ v, const int* __restrict__ stop) {
  if (stop && *stop) return;
  __shared__ double tmp[256];
  const int ncf = NCF ? NCF : ncf_, nbf = NBF ? NBF : nbf_;
  const int fpb = blockDim.x / ncf;
  const int r = threadIdx.x % ncf, fl = threadIdx.x / ncf;
  const int64_t ri = (int64_t)blockIdx.x * fpb + fl;
  const bool active = fl < fpb && ri < nrows;
  const int i = active ? rows[ri] : 0;
  const int n2 = ncf * ncf;
  const double* LUi = LU + (size_t)n2 * nbf * i;
  if (active) {
    const int rk = rank[i];
    double t = y[r + (int64_t)ncf * i];
#pragma unroll
    for (int k = 1; k < nbf; ++k) {
      const int j = f2f[k + (int64_t)nbf * i];
      if (j < 0 || rank[j] < rk) continue;
      const double* U = LUi + (size_t)n2 * k;
      const double* vj = v + (size_t)ncf * j;
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < ncf; ++c) s += U[r + ncf * c] * vj[c];
      t -= s;
    }
    tmp[threadIdx.x] = t;
  }
  __syncthreads();
  if (active) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < ncf; ++c) s += LUi[r + ncf * c] * tmp[fl * ncf + c];
    v[r + (int64_t)ncf * i] = s;
  }
}

// block shapes (ncf, nbf) compiled with fixed sizes: 2D triangles p = 1..4
// with ncu = 4, 3D tetrahedra p = 1, 2 with ncu = 5
#define HDGX_TRACE_SHAPES(X) X(8, 5) X(12, 5) X(16, 5) X(20, 5) X(15, 7) X(30, 7)

// Arnoldi state of one GMRES cycle, kept on the device so the host syncs only
// every few steps: Givens-rotated Hessenberg H (column j at MAXM*j), the
// rotations and the rotated right-hand side g (gmres.m:81-94)
struct GmresDev {
  int stop;           // set when converged, at maxit or on breakdown: later steps are no-ops
  int nj;             // Arnoldi steps completed
  int it;             // gmres.m's iteration count
  int flag;           // -1 running, 0 converged, 1 maxit
  double res;
  double H[MAXM * MAXM];
  double cs[MAXM], sn[MAXM], g[MAXM + 1];
};

__global__ void cycle_init_kernel(GmresDev* d, double beta, int it) {
  if (threadIdx.x != 0) return;
  d->stop = 0;
  d->nj = 0;
  d->it = it;
  d->flag = -1;
  d->res = beta;
  for (int q = 0; q <= MAXM; ++q) d->g[q] = 0.0;
  d->g[0] = beta;
}

// column j of H from the Gram-Schmidt coefficients hcol[0..j] and |w|^2 in
// hcol[j+1]: rotate, extend the rotation, update the residual estimate and
// test the stopping rules that gmres.m checks at the top of its next step
__device__ void arnoldi_tail(GmresDev* d, int j, const double* hcol, double tolb, int maxit) {
  if (d->stop) return;
  double h[MAXM + 1];
  for (int q = 0; q <= j; ++q) h[q] = hcol[q];
  h[j + 1] = sqrt(hcol[j + 1]);
  const bool breakdown = !(h[j + 1] > 0.0);
  for (int q = 0; q < j; ++q) {
    const double a = h[q], b = h[q + 1];
    h[q] = d->cs[q] * a + d->sn[q] * b;
    h[q + 1] = -d->sn[q] * a + d->cs[q] * b;
  }
  const double rr = hypot(h[j], h[j + 1]);
  const double c = rr > 0.0 ? h[j] / rr : 1.0, sv = rr > 0.0 ? h[j + 1] / rr : 0.0;
  d->cs[j] = c;
  d->sn[j] = sv;
  h[j] = rr;
  d->g[j + 1] = -sv * d->g[j];
  d->g[j] = c * d->g[j];
  for (int q = 0; q <= j; ++q) d->H[q + MAXM * j] = h[q];
  d->res = fabs(d->g[j + 1]);
  d->nj = j + 1;
  d->it += 1;
  if (breakdown || d->res <= tolb) { d->stop = 1; d->flag = 0; }
  else if (d->it >= maxit) { d->stop = 1; d->flag = 1; }
}

// ---- reductions: partial[q + m*b] = sum over block b's share of V(:,q) . w
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// out[q] = sum_b partial[q + m*b] by one workgroup of 256 threads, in a fixed
// order (so every run gives the same bits)
__device__ void block_finish(int m, int nb, const double* partial, double* out) {
  __shared__ double red[4][MAXM];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double acc[MAXM];
#pragma unroll
  for (int q = 0; q < MAXM; ++q) acc[q] = 0.0;
  // the partial sums come from an earlier launch on the same stream
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int q = 0; q < MAXM; ++q)
      if (q < m) acc[q] += partial[q + (int64_t)m * b];
#pragma unroll
  for (int q = 0; q < MAXM; ++q)
    if (q < m) {
      const double sq = wave_sum(acc[q]);
      if (lane == 0) red[wv][q] = sq;
    }
  __syncthreads();
  if ((int)threadIdx.x < m) out[threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

__global__ __launch_bounds__(256) void finish_kernel(int m, int nb, const double* __restrict__ partial,
                                                     double* __restrict__ out, const int* __restrict__ stop) {
  if (stop && *stop) return;
  block_finish(m, nb, partial, out);
}

// |w|^2 -> hcol[j+1], then the Givens step of Hessenberg column j
__global__ __launch_bounds__(256) void finish_tail_kernel(int nb, const double* __restrict__ partial, double* hcol,
                                                          GmresDev* d, int j, double tolb, int maxit) {
  if (d->stop) return;
  block_finish(1, nb, partial, hcol + j + 1);
  __syncthreads();
  if (threadIdx.x == 0) arnoldi_tail(d, j, hcol, tolb, maxit);
}

__global__ __launch_bounds__(RED_THREADS) void dots_kernel(int64_t n, int m, const double* __restrict__ V, int64_t ldv,
                                                            const double* __restrict__ w, double* __restrict__ partial,
                                                            const int* __restrict__ stop) {
  if (stop && *stop) return;
  __shared__ double red[RED_THREADS / 64][MAXM];
  double acc[MAXM];
#pragma unroll
  for (int q = 0; q < MAXM; ++q) acc[q] = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const double wt = w[t];
#pragma unroll
    for (int q = 0; q < MAXM; ++q)
      if (q < m) acc[q] += V[t + ldv * q] * wt;
  }
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < MAXM; ++q)
    if (q < m) {
      const double s = wave_sum(acc[q]);
      if (lane == 0) red[wv][q] = s;
    }
  __syncthreads();
  if ((int)threadIdx.x < m) {
    double s = 0.0;
    for (int k = 0; k < RED_THREADS / 64; ++k) s += red[k][threadIdx.x];
    partial[threadIdx.x + (int64_t)m * blockIdx.x] = s;
  }
}

// w -= V(:, 0:m) h  and partial sums of w.w (classical Gram-Schmidt, gmres.m:70-71)
__global__ __launch_bounds__(RED_THREADS) void cgs_update_kernel(int64_t n, int m, const double* __restrict__ V,
                                                                 int64_t ldv, const double* __restrict__ h,
                                                                 double* __restrict__ w, double* __restrict__ partial,
                                                                 const int* __restrict__ stop) {
  if (stop && *stop) return;
  __shared__ double hs[MAXM];
  __shared__ double red[RED_THREADS / 64];
  if ((int)threadIdx.x < m) hs[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  double acc = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < m; ++q) s += V[t + ldv * q] * hs[q];
    const double v = w[t] - s;
    w[t] = v;
    acc += v * v;
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < RED_THREADS / 64; ++k) s += red[k];
    partial[blockIdx.x] = s;
  }
}

// r = b - t and partial sums of r.r
__global__ __launch_bounds__(RED_THREADS) void sub_norm_kernel(int64_t n, const double* __restrict__ b,
                                                               const double* __restrict__ t, double* __restrict__ r,
                                                               double* __restrict__ partial) {
  __shared__ double red[RED_THREADS / 64];
  double acc = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const double v = b[k] - (t ? t[k] : 0.0);
    r[k] = v;
    acc += v * v;
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < RED_THREADS / 64; ++k) s += red[k];
    partial[blockIdx.x] = s;
  }
}

// out = x / sqrt(*nrm2)
__global__ void scale_kernel(int64_t n, const double* __restrict__ x, const double* __restrict__ nrm2,
                             double* __restrict__ out, const int* __restrict__ stop) {
  if (stop && *stop) return;
  const double s = 1.0 / sqrt(*nrm2);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    out[k] = x[k] * s;
}

// x += V(:, 0:m) y
__global__ void gemv_add_kernel(int64_t n, int m, const double* __restrict__ V, int64_t ldv,
                                const double* __restrict__ y, double* __restrict__ x) {
  __shared__ double ys[MAXM];
  if ((int)threadIdx.x < m) ys[threadIdx.x] = y[threadIdx.x];
  __syncthreads();
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < m; ++q) s += V[k + ldv * q] * ys[q];
    x[k] += s;
  }
}

// ---- Newton update pieces
__global__ void gather_kernel(int ncu, int ndf, int64_t ne, const int32_t* __restrict__ elcon,
                              const double* __restrict__ UH, double* __restrict__ UHE) {
  const int64_t n = (int64_t)ncu * ndf * ne;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(t % ncu);
    const int64_t je = t / ncu;
    UHE[t] = UH[c + (int64_t)ncu * elcon[je]];
  }
}

// RH(c, m, i) = sum over the one or two elements of face i of their Rh at the
// local face node that elcon maps to trace node i*npf + m (the sparse sum of
// hdg_residual.m:41-52), one thread per trace unknown, no atomics
__global__ void trace_scatter_kernel(int ncu, int npf, int nfe, int64_t nf, const int32_t* __restrict__ fconn,
                                     const int32_t* __restrict__ elcon, const double* __restrict__ Rhe,
                                     double* __restrict__ RH) {
  const int64_t n = (int64_t)ncu * npf * nf;
  const int ndf = npf * nfe;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(t % ncu);
    const int64_t node = t / ncu;               // i*npf + m
    const int64_t i = node / npf;
    double v = 0.0;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int64_t e = fconn[2 * side + 4 * i];
      if (e < 0) continue;
      const int j = fconn[2 * side + 1 + 4 * i];
      for (int a = 0; a < npf; ++a)
        if (elcon[a + npf * j + (int64_t)ndf * e] == node) v += Rhe[c + ncu * (a + npf * j) + (int64_t)ncu * ndf * e];
    }
    RH[t] = v;
  }
}

// DUDGT(:, e) = DUDG(:, e) + DUDG_DUH(:, :, e) DUH(:, elcon(:, e)), one element per workgroup
__global__ __launch_bounds__(256) void local_update_kernel(int nr, int ncu, int ndf, int64_t ne,
                                                           const int32_t* __restrict__ elcon,
                                                           const double* __restrict__ DUDG,
                                                           const double* __restrict__ DUDG_DUH,
                                                           const double* __restrict__ DUH,
                                                           double* __restrict__ DUDGT) {
  extern __shared__ double sU[];
  const int64_t e = blockIdx.x;
  if (e >= ne) return;
  const int nh = ncu * ndf;
  for (int t = threadIdx.x; t < nh; t += blockDim.x)
    sU[t] = DUH[t % ncu + (int64_t)ncu * elcon[t / ncu + (int64_t)ndf * e]];
  __syncthreads();
  const double* X = DUDG_DUH + (size_t)nr * nh * e;
  for (int row = threadIdx.x; row < nr; row += blockDim.x) {
    double v = 0.0;
    for (int col = 0; col < nh; ++col) v += X[row + (size_t)nr * col] * sU[col];
    DUDGT[row + (int64_t)nr * e] = DUDG[row + (int64_t)nr * e] + v;
  }
}

__global__ void newton_update_kernel(int npv, int nc, int nco, int64_t ne, int64_t nuh, double alpha,
                                     const double* UDG0, const double* __restrict__ DUDGT, const double* UH0,
                                     const double* __restrict__ DUH, double* UDG, double* UH) {
  const int64_t nu = (int64_t)npv * nc * ne;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nu; t += stride) {
    const int64_t e = t / ((int64_t)npv * nc);
    const int rc = (int)(t % ((int64_t)npv * nc));
    if (rc < npv * nco) UDG[t] = UDG0[t] + alpha * DUDGT[rc + (int64_t)npv * nco * e];
    else if (UDG != UDG0) UDG[t] = UDG0[t];
  }
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nuh; t += stride) UH[t] = UH0[t] + alpha * DUH[t];
}

unsigned grid_for(int64_t n, int threads, unsigned cap = 4096) {
  const int64_t g = (n + threads - 1) / threads;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

int check_trace(hdgx_ctx* ctx, const hdgx_trace* T) {
  if (!T) return fail(ctx, HDGX_ERR_ARG, "trace: null descriptor");
  if (T->ncu < 1 || T->npf < 1 || T->nfe < 2 || T->nf < 0 || T->ne < 0) return fail(ctx, HDGX_ERR_ARG, "trace: sizes");
  if (T->ncu * T->npf > 64) return fail(ctx, HDGX_ERR_UNSUPPORTED, "trace: ncu*npf > 64");
  if (T->nfe > 4) return fail(ctx, HDGX_ERR_UNSUPPORTED, "trace: nfe > 4");
  if (T->nf > 0 && (!T->fconn || !T->f2f)) return fail(ctx, HDGX_ERR_ARG, "trace: null connectivity");
  return HDGX_OK;
}

int check_colors(hdgx_ctx* ctx, const hdgx_trace* T) {
  if (T->ncolors < 1 || !T->color_ptr || !T->color_rows || !T->face_rank)
    return fail(ctx, HDGX_ERR_ARG, "trace: block ILU(0) needs the colouring");
  if (T->color_ptr[0] != 0 || T->color_ptr[T->ncolors] != T->nf) return fail(ctx, HDGX_ERR_ARG, "trace: color_ptr");
  for (int c = 0; c < T->ncolors; ++c)
    if (T->color_ptr[c + 1] < T->color_ptr[c]) return fail(ctx, HDGX_ERR_ARG, "trace: color_ptr not monotone");
  return HDGX_OK;
}

int ensure_tws(hdgx_ctx* ctx, size_t bytes) {
  if (!ctx->pinned) HIP_TRY(ctx, hipHostMalloc((void**)&ctx->pinned, 2048 * sizeof(double)));
  if (ctx->tws_bytes >= bytes) return HDGX_OK;
  if (ctx->tws) HIP_TRY(ctx, hipFree(ctx->tws));
  ctx->tws = nullptr;
  ctx->tws_bytes = 0;
  HIP_TRY(ctx, hipMalloc(&ctx->tws, bytes));
  ctx->tws_bytes = bytes;
  return HDGX_OK;
}

int launch_matvec(hdgx_ctx* ctx, const hdgx_trace* T, const double* K, const double* x, double* y,
                  const int* stop = nullptr, const double* nrm2 = nullptr, double* vout = nullptr) {
  const int ncf = T->ncu * T->npf, nbf = 2 * T->nfe - 1;
  const int fpb = 256 / ncf;
  const unsigned grid = (unsigned)((T->nf + fpb - 1) / fpb);
  if (grid) {
    bool done = false;
#define HDGX_MV(NC, NB)                                                                                           \
    if (!done && ncf == NC && nbf == NB) {                                                                         \
      hipLaunchKernelGGL((trace_matvec_kernel<NC, NB>), dim3(grid), dim3(fpb * ncf), 0, ctx->stream, ncf, nbf,   \
                         T->nf, K, T->f2f, x, y, nrm2, vout, stop);                                                \
      done = true;                                                                                                 \
    }
    HDGX_TRACE_SHAPES(HDGX_MV)
#undef HDGX_MV
    if (!done)
      hipLaunchKernelGGL((trace_matvec_kernel<0, 0>), dim3(grid), dim3(fpb * ncf), 0, ctx->stream, ncf, nbf, T->nf, K,
                         T->f2f, x, y, nrm2, vout, stop);
  }
  HIP_TRY(ctx, hipGetLastError());
  return HDGX_OK;
}

int launch_bilu0_apply(hdgx_ctx* ctx, const hdgx_trace* T, const double* LU, const double* x, double* y, double* tmp,
                       const int* stop = nullptr) {
  const int ncf = T->ncu * T->npf, nbf = 2 * T->nfe - 1;
  const int fpb = 256 / ncf;
  const int C = T->ncolors;
  if (C < 1) return HDGX_OK;
  const int n0 = T->color_ptr[1] - T->color_ptr[0];
  // lower sweep of colour c (the last one also applies U_ii^-1 into y) or
  // upper sweep of colour c reading the lower solution from ysrc
  auto sweep = [&](bool lower, int c, bool last, const double* ysrc) {
    const int n = T->color_ptr[c + 1] - T->color_ptr[c];
    if (n == 0) return;
    const int32_t* rows = T->color_rows + T->color_ptr[c];
    const dim3 grid((n + fpb - 1) / fpb), block(fpb * ncf);
    double* vd = last ? y : nullptr;
#define HDGX_SW(NC, NB)                                                                                           \
    if (ncf == NC && nbf == NB) {                                                                                  \
      if (lower)                                                                                                   \
        hipLaunchKernelGGL((bilu0_lower_kernel<NC, NB>), grid, block, 0, ctx->stream, ncf, nbf, rows, n, T->f2f,  \
                           T->face_rank, LU, x, tmp, n0, vd, stop);                                                \
      else                                                                                                         \
        hipLaunchKernelGGL((bilu0_upper_kernel<NC, NB>), grid, block, 0, ctx->stream, ncf, nbf, rows, n, T->f2f,  \
                           T->face_rank, LU, ysrc, y, stop);                                                       \
      return;                                                                                                      \
    }
    HDGX_TRACE_SHAPES(HDGX_SW)
#undef HDGX_SW
    if (lower)
      hipLaunchKernelGGL((bilu0_lower_kernel<0, 0>), grid, block, 0, ctx->stream, ncf, nbf, rows, n, T->f2f,
                         T->face_rank, LU, x, tmp, n0, vd, stop);
    else
      hipLaunchKernelGGL((bilu0_upper_kernel<0, 0>), grid, block, 0, ctx->stream, ncf, nbf, rows, n, T->f2f,
                         T->face_rank, LU, ysrc, y, stop);
  };
  // the first colour's lower solve is the identity (no earlier neighbours):
  // later colours read its y from x; the last colour's upper solve is its
  // diagonal block alone and runs inside its lower sweep
  for (int c = C > 1 ? 1 : 0; c < C; ++c) sweep(true, c, c == C - 1, nullptr);
  for (int c = C - 2; c >= 0; --c) sweep(false, c, false, c == 0 ? x : tmp);
  HIP_TRY(ctx, hipGetLastError());
  return HDGX_OK;
}

// sum of m values reduced from the partial buffer into dev_out, copied to host
int reduce_to_host(hdgx_ctx* ctx, int m, int nb, const double* partial, double* dev_out, double* host) {
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(256), 0, ctx->stream, m, nb, partial, dev_out, (const int*)nullptr);
  HIP_TRY(ctx, hipGetLastError());
  if (host) {
    HIP_TRY(ctx, hipMemcpyAsync(ctx->pinned, dev_out, m * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (int q = 0; q < m; ++q) host[q] = ctx->pinned[q];
  }
  return HDGX_OK;
}

}  // namespace

extern "C" {

int hdgx_trace_assemble(hdgx_ctx* ctx, const hdgx_trace* T, const double* AE, const double* FE, double* K, double* F,
                        int32_t* status) {
  if (!ctx) return HDGX_ERR_ARG;
  int rc = check_trace(ctx, T);
  if (rc) return rc;
  if (T->nf == 0) return HDGX_OK;
  if (!T->elcon || !AE || !FE || !K || !F) return fail(ctx, HDGX_ERR_ARG, "trace_assemble: null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int ncf = T->ncu * T->npf, nbf = 2 * T->nfe - 1;
  const size_t lds = ((size_t)ncf * (ncf * nbf + 1) + ncf) * sizeof(double) + (size_t)(2 + nbf) * T->npf * 4 + 16;
  if (lds > 64 * 1024) return fail(ctx, HDGX_ERR_UNSUPPORTED, "trace_assemble: face block too large");
  hipLaunchKernelGGL(trace_assemble_kernel, dim3((unsigned)T->nf), dim3(64), lds, ctx->stream, T->ncu, T->npf, T->nfe,
                     T->nf, T->fconn, T->f2f, T->elcon, AE, FE, K, F, status);
  HIP_TRY(ctx, hipGetLastError());
  return HDGX_OK;
}

int hdgx_trace_matvec(hdgx_ctx* ctx, const hdgx_trace* T, const double* K, const double* x, double* y) {
  if (!ctx) return HDGX_ERR_ARG;
  int rc = check_trace(ctx, T);
  if (rc) return rc;
  if (T->nf == 0) return HDGX_OK;
  if (!K || !x || !y) return fail(ctx, HDGX_ERR_ARG, "trace_matvec: null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  return launch_matvec(ctx, T, K, x, y);
}

int hdgx_bilu0_factor(hdgx_ctx* ctx, const hdgx_trace* T, const double* K, double* LU, int32_t* status) {
  if (!ctx) return HDGX_ERR_ARG;
  int rc = check_trace(ctx, T);
  if (rc) return rc;
  if (T->nf == 0) return HDGX_OK;
  if ((rc = check_colors(ctx, T))) return rc;
  if (!K || !LU) return fail(ctx, HDGX_ERR_ARG, "bilu0_factor: null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int ncf = T->ncu * T->npf, nbf = 2 * T->nfe - 1;
  const size_t lds = ((size_t)ncf * ncf * (nbf + 2) + ncf) * sizeof(double) + (nbf + 1) * 4 + 16;
  if (lds > 64 * 1024) return fail(ctx, HDGX_ERR_UNSUPPORTED, "bilu0_factor: face block too large");
  for (int c = 0; c < T->ncolors; ++c) {
    const int n = T->color_ptr[c + 1] - T->color_ptr[c];
    if (n == 0) continue;
    hipLaunchKernelGGL(bilu0_row_kernel, dim3(n), dim3(64), lds, ctx->stream, ncf, nbf, T->color_rows + T->color_ptr[c],
                       n, T->f2f, T->face_rank, K, LU, status);
  }
  HIP_TRY(ctx, hipGetLastError());
  return HDGX_OK;
}

int hdgx_bilu0_apply(hdgx_ctx* ctx, const hdgx_trace* T, const double* LU, const double* x, double* y) {
  if (!ctx) return HDGX_ERR_ARG;
  int rc = check_trace(ctx, T);
  if (rc) return rc;
  if (T->nf == 0) return HDGX_OK;
  if ((rc = check_colors(ctx, T))) return rc;
  if (!LU || !x || !y || x == y) return fail(ctx, HDGX_ERR_ARG, "bilu0_apply: null or aliased array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t N = (size_t)T->ncu * T->npf * T->nf;
  if ((rc = ensure_tws(ctx, N * sizeof(double)))) return rc;
  return launch_bilu0_apply(ctx, T, LU, x, y, static_cast<double*>(ctx->tws));
}

int hdgx_gmres(hdgx_ctx* ctx, const hdgx_trace* T, const double* K, const double* LU, const double* F, double* x,
               int restart, double tol, int maxit, int* iters, double* relres, int* flag) {
  if (!ctx) return HDGX_ERR_ARG;
  int rc = check_trace(ctx, T);
  if (rc) return rc;
  if (restart < 1 || restart + 1 > MAXM) return fail(ctx, HDGX_ERR_ARG, "gmres: restart must be 1..31");
  if (T->nf == 0) {
    if (iters) *iters = 0;
    if (relres) *relres = 0.0;
    if (flag) *flag = 0;
    return HDGX_OK;
  }
  if (!K || !F || !x) return fail(ctx, HDGX_ERR_ARG, "gmres: null array");
  if (LU && (rc = check_colors(ctx, T))) return rc;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int64_t N = (int64_t)T->ncu * T->npf * T->nf;
  if (maxit <= 0) maxit = (int)std::min<int64_t>(N, 1 << 30);
  const int m1 = restart + 1;
  // workspace: V (N x m1) | t | u | b~ | w0 | w1 | partial | h | y | Arnoldi state
  const size_t dwords = (sizeof(GmresDev) + 7) / 8;
  const size_t words = (size_t)N * (m1 + 5) + (size_t)RED_BLOCKS * MAXM + 2 * MAXM + dwords;
  if ((rc = ensure_tws(ctx, words * sizeof(double)))) return rc;
  double* V = static_cast<double*>(ctx->tws);
  double* t = V + (size_t)N * m1;
  double* u = t + N;
  double* bt = u + N;
  double* wk[2] = {bt + N, bt + 2 * N};     // Krylov vectors before normalisation
  double* partial = bt + 3 * N;
  double* hdev = partial + (size_t)RED_BLOCKS * MAXM;
  double* ydev = hdev + MAXM;
  GmresDev* dst = reinterpret_cast<GmresDev*>(ydev + MAXM);
  const int* stop = &dst->stop;
  const unsigned rg = grid_for(N, RED_THREADS, RED_BLOCKS);
  hipStream_t s = ctx->stream;
  constexpr int SYNC_EVERY = 4;     // Arnoldi steps launched between host syncs

  // operator: out = A v (plus the ILU(0) solve when LU is given); with nrm2
  // the operand is v / sqrt(*nrm2), written normalised to vout
  auto op = [&](const double* v, double* out, const int* stp, const double* nrm2 = nullptr,
                double* vout = nullptr) -> int {
    if (!LU) return launch_matvec(ctx, T, K, v, out, stp, nrm2, vout);
    int r = launch_matvec(ctx, T, K, v, u, stp, nrm2, vout);
    if (r) return r;
    return launch_bilu0_apply(ctx, T, LU, u, out, t, stp);   // out is never t
  };

  // b~ = P^-1 F (gmres_bilu0.m:18) or F
  const double* b = F;
  if (LU) {
    if ((rc = launch_bilu0_apply(ctx, T, LU, F, bt, t))) return rc;
    b = bt;
  }
  double h[MAXM];
  hipLaunchKernelGGL(sub_norm_kernel, dim3(rg), dim3(RED_THREADS), 0, s, N, b, (const double*)nullptr, t, partial);
  if ((rc = reduce_to_host(ctx, 1, rg, partial, hdev, h))) return rc;
  const double nrmb = std::sqrt(h[0]);
  int it = 0, fl = -1;
  double res = 0.0;
  if (!(nrmb > 0.0)) {
    HIP_TRY(ctx, hipMemsetAsync(x, 0, N * sizeof(double), s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (iters) *iters = 0;
    if (relres) *relres = 0.0;
    if (flag) *flag = 0;
    return HDGX_OK;
  }
  const double tolb = tol * nrmb;
  GmresDev* hst = reinterpret_cast<GmresDev*>(ctx->pinned);
  while (true) {
    // r = b - A x (gmres.m:37-47)
    if ((rc = op(x, V, nullptr))) return rc;                 // V(:,0) as scratch for A x
    hipLaunchKernelGGL(sub_norm_kernel, dim3(rg), dim3(RED_THREADS), 0, s, N, b, (const double*)V, t, partial);
    if ((rc = reduce_to_host(ctx, 1, rg, partial, hdev, h))) return rc;
    const double beta = std::sqrt(h[0]);
    hipLaunchKernelGGL(scale_kernel, dim3(rg), dim3(RED_THREADS), 0, s, N, (const double*)t, (const double*)hdev, V,
                       (const int*)nullptr);
    res = beta;
    ++it;
    int nj = 0;
    if (res <= tolb) fl = 0;
    else if (it >= maxit) fl = 1;
    if (fl < 0) {
      hipLaunchKernelGGL(cycle_init_kernel, dim3(1), dim3(64), 0, s, dst, beta, it);
      for (int j = 0; j < restart; ++j) {
        // step j's operand: V(:,0), or the previous step's w, normalised by
        // the matvec into V(:,j) with |w|^2 from hcol[j]
        double* w = wk[j & 1];
        if (j == 0) rc = op(V, w, stop);
        else rc = op(wk[(j - 1) & 1], w, stop, hdev + j, V + (size_t)N * j);
        if (rc) return rc;
        // h(0:j) = V(:,0:j)' w ; w -= V h ; |w|^2 (gmres.m:80-83)
        hipLaunchKernelGGL(dots_kernel, dim3(rg), dim3(RED_THREADS), 0, s, N, j + 1, (const double*)V, N,
                           (const double*)w, partial, stop);
        hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(256), 0, s, j + 1, (int)rg, (const double*)partial, hdev, stop);
        hipLaunchKernelGGL(cgs_update_kernel, dim3(rg), dim3(RED_THREADS), 0, s, N, j + 1, (const double*)V, N,
                           (const double*)hdev, w, partial + (size_t)RED_BLOCKS * (MAXM - 1), stop);
        hipLaunchKernelGGL(finish_tail_kernel, dim3(1), dim3(256), 0, s, (int)rg,
                           (const double*)(partial + (size_t)RED_BLOCKS * (MAXM - 1)), hdev, dst, j, tolb, maxit);
        HIP_TRY(ctx, hipGetLastError());
        if ((j + 1) % SYNC_EVERY == 0 || j == restart - 1) {
          HIP_TRY(ctx, hipMemcpyAsync(hst, dst, offsetof(GmresDev, H), hipMemcpyDeviceToHost, s));
          HIP_TRY(ctx, hipStreamSynchronize(s));
          if (hst->stop) break;
        }
      }
      HIP_TRY(ctx, hipMemcpyAsync(hst, dst, sizeof(GmresDev), hipMemcpyDeviceToHost, s));
      HIP_TRY(ctx, hipStreamSynchronize(s));
      nj = hst->nj;
      it = hst->it;
      res = hst->res;
      fl = hst->flag;
      // x += V(:, 0:nj) y with R y = g (gmres.m:101)
      if (nj > 0) {
        double y[MAXM];
        for (int q = nj - 1; q >= 0; --q) {
          double v = hst->g[q];
          for (int k = q + 1; k < nj; ++k) v -= hst->H[q + MAXM * k] * y[k];
          y[q] = v / hst->H[q + MAXM * q];
        }
        double* ys = ctx->pinned + (sizeof(GmresDev) + 7) / 8;
        for (int q = 0; q < nj; ++q) ys[q] = y[q];
        HIP_TRY(ctx, hipMemcpyAsync(ydev, ys, nj * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(gemv_add_kernel, dim3(grid_for(N, 256)), dim3(256), 0, s, N, nj, (const double*)V, N,
                           (const double*)ydev, x);
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, hipStreamSynchronize(s));
      }
    }
    if (fl >= 0) break;
  }
  if (iters) *iters = it;
  if (relres) *relres = res / nrmb;
  if (flag) *flag = fl;
  return HDGX_OK;
}

int hdgx_color_faces(int nbf, int64_t nf, const int32_t* f2f, int32_t* color, int32_t* ncolors) {
  if (nbf < 1 || nf < 0 || !ncolors || (nf > 0 && (!f2f || !color))) return HDGX_ERR_ARG;
  // greedy in face order: the smallest colour none of the already coloured
  // neighbours has (a face has nbf - 1 neighbours, so nbf colours suffice)
  std::vector<char> used(nbf + 1);
  int nc = 0;
  for (int64_t i = 0; i < nf; ++i) color[i] = -1;
  for (int64_t i = 0; i < nf; ++i) {
    std::fill(used.begin(), used.end(), 0);
    for (int k = 1; k < nbf; ++k) {
      const int32_t j = f2f[k + (int64_t)nbf * i];
      if (j >= nf) return HDGX_ERR_ARG;
      if (j >= 0 && color[j] >= 0) used[color[j]] = 1;
    }
    int c = 0;
    while (used[c]) ++c;
    color[i] = c;
    nc = std::max(nc, c + 1);
  }
  *ncolors = nc;
  return HDGX_OK;
}

int hdgx_trace_gather(hdgx_ctx* ctx, int ncu, int npf, int nfe, int64_t ne, const int32_t* elcon, const double* UH,
                      double* UHE) {
  if (!ctx) return HDGX_ERR_ARG;
  if (ne < 0 || ncu < 1 || npf < 1 || nfe < 1) return fail(ctx, HDGX_ERR_ARG, "trace_gather: sizes");
  if (ne == 0) return HDGX_OK;
  if (!elcon || !UH || !UHE) return fail(ctx, HDGX_ERR_ARG, "trace_gather: null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int64_t n = (int64_t)ncu * npf * nfe * ne;
  hipLaunchKernelGGL(gather_kernel, dim3(grid_for(n, 256, 65536)), dim3(256), 0, ctx->stream, ncu, npf * nfe, ne, elcon,
                     UH, UHE);
  HIP_TRY(ctx, hipGetLastError());
  return HDGX_OK;
}

int hdgx_trace_scatter(hdgx_ctx* ctx, const hdgx_trace* T, const double* Rhe, double* RH) {
  if (!ctx) return HDGX_ERR_ARG;
  int rc = check_trace(ctx, T);
  if (rc) return rc;
  if (T->nf == 0) return HDGX_OK;
  if (!T->elcon || !Rhe || !RH) return fail(ctx, HDGX_ERR_ARG, "trace_scatter: null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int64_t n = (int64_t)T->ncu * T->npf * T->nf;
  hipLaunchKernelGGL(trace_scatter_kernel, dim3(grid_for(n, 256, 65536)), dim3(256), 0, ctx->stream, T->ncu, T->npf,
                     T->nfe, T->nf, T->fconn, T->elcon, Rhe, RH);
  HIP_TRY(ctx, hipGetLastError());
  return HDGX_OK;
}

int hdgx_local_update(hdgx_ctx* ctx, int npv, int nco, int ncu, int npf, int nfe, int64_t ne, const int32_t* elcon,
                      const double* DUDG, const double* DUDG_DUH, const double* DUH, double* DUDGT) {
  if (!ctx) return HDGX_ERR_ARG;
  if (ne < 0 || npv < 1 || nco < 1 || ncu < 1 || npf < 1 || nfe < 1) return fail(ctx, HDGX_ERR_ARG, "local_update: sizes");
  if (ne == 0) return HDGX_OK;
  if (!elcon || !DUDG || !DUDG_DUH || !DUH || !DUDGT) return fail(ctx, HDGX_ERR_ARG, "local_update: null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int nr = npv * nco, ndf = npf * nfe;
  const int threads = std::min(256, ((nr + 63) / 64) * 64);
  const size_t lds = (size_t)ncu * ndf * sizeof(double);
  if (ne > 0x7fffffff) return fail(ctx, HDGX_ERR_ARG, "local_update: too many elements for one launch");
  hipLaunchKernelGGL(local_update_kernel, dim3((unsigned)ne), dim3(threads), lds, ctx->stream, nr, ncu, ndf, ne, elcon,
                     DUDG, DUDG_DUH, DUH, DUDGT);
  HIP_TRY(ctx, hipGetLastError());
  return HDGX_OK;
}

int hdgx_newton_update(hdgx_ctx* ctx, int npv, int nc, int nco, int64_t ne, int64_t nuh, double alpha,
                       const double* UDG0, const double* DUDGT, const double* UH0, const double* DUH, double* UDG,
                       double* UH) {
  if (!ctx) return HDGX_ERR_ARG;
  if (ne < 0 || nuh < 0 || npv < 1 || nco < 1 || nco > nc) return fail(ctx, HDGX_ERR_ARG, "newton_update: sizes");
  if ((ne > 0 && (!UDG0 || !DUDGT || !UDG)) || (nuh > 0 && (!UH0 || !DUH || !UH)))
    return fail(ctx, HDGX_ERR_ARG, "newton_update: null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int64_t n = std::max<int64_t>((int64_t)npv * nc * ne, nuh);
  if (n == 0) return HDGX_OK;
  hipLaunchKernelGGL(newton_update_kernel, dim3(grid_for(n, 256, 65536)), dim3(256), 0, ctx->stream, npv, nc, nco, ne,
                     nuh, alpha, UDG0, DUDGT, UH0, DUH, UDG, UH);
  HIP_TRY(ctx, hipGetLastError());
  return HDGX_OK;
}

int hdgx_norm2(hdgx_ctx* ctx, int64_t n, const double* x, double* result) {
  if (!ctx) return HDGX_ERR_ARG;
  if (n < 0 || !result || (n > 0 && !x)) return fail(ctx, HDGX_ERR_ARG, "norm2: arguments");
  if (n == 0) { *result = 0.0; return HDGX_OK; }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = ensure_tws(ctx, ((size_t)RED_BLOCKS + 2) * sizeof(double));
  if (rc) return rc;
  // the trace workspace may hold a Krylov basis; the reduction uses its tail
  double* partial = static_cast<double*>(ctx->tws) + (ctx->tws_bytes / sizeof(double) - RED_BLOCKS - 2);
  const unsigned rg = grid_for(n, RED_THREADS, RED_BLOCKS);
  hipLaunchKernelGGL(dots_kernel, dim3(rg), dim3(RED_THREADS), 0, ctx->stream, n, 1, x, n, x, partial, (const int*)nullptr);
  double h[1];
  if ((rc = reduce_to_host(ctx, 1, rg, partial, partial + RED_BLOCKS, h))) return rc;
  *result = std::sqrt(h[0]);
  return HDGX_OK;
}

}  // extern "C"
