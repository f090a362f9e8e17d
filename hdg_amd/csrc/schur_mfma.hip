// FP64 MFMA form of the batched Schur elimination (kernel/schur_u.c:42-163,
// kernel/schur_uq.c:40-287) for gfx950.
//
// One wavefront per element.  The element system S = [D' | RHS] (D' = D or
// D + B M^-1 C; RHS = [Ru, -F] or [Ru - B M^-1 Rq, B M^-1 E - F]) lives in
// v_mfma_f64_16x16x4f64 accumulator tiles: lane l = 16g + c, register i of a
// tile holds row 4i + g, column c (4g + i in the FP32 form's tiles).  Gauss-Jordan elimination with partial
// pivoting runs in panels of four columns: the panel is factored one row per
// lane (wavefront argmax for the pivot), which yields the rank-4 transform
// S <- S + Z * S[pivot rows, :]; that update is applied to every tile right
// of the panel with MFMA.  After the last panel row p_k of the RHS tiles holds
// solution row k.  B M^-1 C and B M^-1 E are MFMA products whose all-zero
// k-steps are skipped at compile time, and AE = H + GK X runs per AE row tile
// over only the columns of GK that touch that tile's faces (GK is face
// sparse, faceint.m:354-358).
//
// RHS columns are stored in the order q = 0 (du), q = 1 + jf + NDF*k for the
// trace unknown (k, jf), so each 16-column tile meets few state components.
#include <hip/hip_runtime.h>

#include "hdgx_device.h"
#include "hdgx_kernels.h"

namespace hdgx {

// the 16x16x4 MFMA of the elimination's element type.  Both forms take the A
// operand as row l%16, k l/16 and B as k l/16, column l%16 (lane l = 16g + c);
// their accumulators differ: register i holds row 4i + g (FP64) or row
// 4g + i (FP32), column c.  row(i, g) is that row; reg / grp invert it.
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef d4 V;
  typedef unsigned long long K;   // pivot search key
  static HDGX_DEV d4 op(double a, double b, d4 c) { return mfma64(a, b, c); }
  static HDGX_DEV constexpr int row(int i, int g) { return 4 * i + g; }
  static HDGX_DEV constexpr int reg(int r) { return r >> 2; }
  static HDGX_DEV constexpr int grp(int r) { return r & 3; }
  static HDGX_DEV K key(double a) {
    return isnan(a) ? 0x7ff0000000000000ull : (unsigned long long)__double_as_longlong(a);
  }
};
template <> struct Mfma<float> {
  typedef f4 V;
  typedef unsigned K;
  static HDGX_DEV f4 op(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static HDGX_DEV constexpr int row(int i, int g) { return 4 * g + i; }
  static HDGX_DEV constexpr int reg(int r) { return r & 3; }
  static HDGX_DEV constexpr int grp(int r) { return r >> 2; }
  static HDGX_DEV K key(float a) { return isnan(a) ? 0x7f800000u : __float_as_uint(a); }
};

// most FP64 refinement sweeps of the FP32 form
constexpr int IR_SWEEPS = 4;
// no column: entry of the AE k-step table past an AE row tile's node set
constexpr int AE_NONE = 255;

// value of lane `src` (wavefront-uniform) in every lane, through v_readlane
HDGX_DEV double lane_bcast(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}
HDGX_DEV float lane_bcast(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// maximum of a pivot-search key over the wavefront, wavefront-uniform: DPP
// rotations within each row of 16 lanes, then the four row maxima read out
template <int SH>
HDGX_DEV unsigned row_ror(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x120 + SH, 0xf, 0xf, false);
}
template <int SH>
HDGX_DEV unsigned long long key_max_ror(unsigned long long k) {
  const unsigned long long o = ((unsigned long long)row_ror<SH>((unsigned)(k >> 32)) << 32) | row_ror<SH>((unsigned)k);
  return o > k ? o : k;
}
template <int SH>
HDGX_DEV unsigned key_max_ror(unsigned k) {
  const unsigned o = row_ror<SH>(k);
  return o > k ? o : k;
}
template <typename K>
HDGX_DEV K wave_key_max(K k) {
  k = key_max_ror<8>(k);
  k = key_max_ror<4>(k);
  k = key_max_ror<2>(k);
  k = key_max_ror<1>(k);
  K m = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    K v;
    if constexpr (sizeof(K) == 8)
      v = ((K)(unsigned)__builtin_amdgcn_readlane((int)(k >> 32), 16 * r) << 32) |
          (unsigned)__builtin_amdgcn_readlane((int)k, 16 * r);
    else
      v = (unsigned)__builtin_amdgcn_readlane((int)k, 16 * r);
    m = v > m ? v : m;
  }
  return m;
}

// FP64 pivot-search key of |a| in lane `lane`: the exponent and leading 15
// mantissa bits of |a| above the low six bits 63 - lane, so one unsigned
// wavefront max finds the largest magnitude (to 2^-15 relative; rows closer
// than that count as tied) and ties go to the lowest lane.  NaN counts as
// infinite.
HDGX_DEV unsigned piv_key(double a, int lane) {
  const unsigned long long b = isnan(a) ? 0x7ff0000000000000ull : (unsigned long long)__double_as_longlong(fabs(a));
  return ((unsigned)(b >> 31) & ~63u) | (unsigned)(63 - lane);
}

// 1/x to about an ulp: v_rcp_f64 and two Newton steps, a shorter dependent
// chain than the correctly rounded division
HDGX_DEV double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// maximum over the wavefront
HDGX_DEV double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

template <class C>
struct MfmaShape {
  static constexpr int NU = C::NU, NH = C::NH, NRHS = 1 + C::NH;
  static constexpr int RT = (NU + 15) / 16;            // row tiles (= D' column tiles)
  static constexpr int R = 16 * RT;
  static constexpr int TRALL = (NRHS + 15) / 16;       // RHS tiles
  // RHS tiles per pass: keep the accumulators at <= 96 (80 for 4 row tiles)
  // doubles per lane
  static constexpr int ACCB = RT >= 4 ? 80 : 96;
  static constexpr int TRB0 = (ACCB / (4 * RT)) - RT;
  static constexpr int TRB = TRB0 < 1 ? 1 : (TRB0 > TRALL ? TRALL : TRB0);
  static constexpr int NPASS = (TRALL + TRB - 1) / TRB;
  static constexpr int CTP = RT + TRB;                 // tiles per pass
  static constexpr int NPAN = (NU + 3) / 4;
  static constexpr int RTH = (NH + 15) / 16;           // AE row tiles
  static constexpr int KB = C::NPV * C::NCU * C::ND;   // columns of B (q part of BD)
  static constexpr int KBS = (KB + 3) / 4;
  // wavefronts per SIMD the register budget is set for: 4 / 3 / 2 for one /
  // two / three row tiles (accumulators in VGPRs, no spills), one beyond
  // (accumulators in AGPRs)
  static constexpr int WPE = RT == 1 ? 4 : (RT == 2 ? 3 : (RT == 3 ? 2 : 1));
  // the FP32 + refinement form: its LDS (every panel transform kept) allows
  // about two wavefronts per SIMD from two row tiles on
  static constexpr int WPE_IR = RT == 1 ? 4 : (RT <= 3 ? 2 : 1);
};

// state component of RHS column q >= 1 and its face-node index jf
template <class C> HDGX_DEV constexpr int rhs_k(int q) { return (q - 1) / C::NDF; }
template <class C> HDGX_DEV constexpr int rhs_jf(int q) { return (q - 1) % C::NDF; }
// original trace column (k + NCU*jf) of RHS column q >= 1
template <class C> HDGX_DEV constexpr int rhs_t(int q) { return rhs_k<C>(q) + C::NCU * rhs_jf<C>(q); }

// does k-step s (B columns 4s..4s+3, B column K = a + NPV*m + NPV*NCU*d) meet
// component m for some column of the D' tile ct (D' column b + NPV*m)?
template <class C>
HDGX_DEV constexpr bool bmic_needed(int ct, int s) {
  for (int kk = 4 * s; kk < 4 * s + 4 && kk < MfmaShape<C>::KB; ++kk) {
    const int m = (kk / C::NPV) % C::NCU;
    for (int col = 16 * ct; col < 16 * ct + 16 && col < C::NU; ++col)
      if (col / C::NPV == m) return true;
  }
  return false;
}

// the same for the RHS tile with global RHS index base q0 (columns q0..q0+15)
template <class C>
HDGX_DEV constexpr bool bmie_needed(int q0, int s) {
  for (int kk = 4 * s; kk < 4 * s + 4 && kk < MfmaShape<C>::KB; ++kk) {
    const int m = (kk / C::NPV) % C::NCU;
    for (int q = q0; q < q0 + 16 && q < 1 + C::NH; ++q)
      if (q >= 1 && rhs_k<C>(q) == m) return true;
  }
  return false;
}

// bit s set: k-step s of B M^-1 C (tile ct) / of B M^-1 E (pass, RHS tile tb)
// meets a nonzero block; the first RHS tile also carries -B M^-1 Rq in column 0
template <class C>
struct SkipMasks {
  uint64_t ct[MfmaShape<C>::RT];
  uint64_t rhs[MfmaShape<C>::NPASS][MfmaShape<C>::TRB];
};

template <class C>
HDGX_DEV constexpr SkipMasks<C> skip_masks() {
  using S = MfmaShape<C>;
  static_assert(S::KBS <= 64, "k-steps of B fit a 64-bit mask");
  SkipMasks<C> m{};
  for (int ct = 0; ct < S::RT; ++ct)
    for (int s = 0; s < S::KBS; ++s)
      if (bmic_needed<C>(ct, s)) m.ct[ct] |= 1ull << s;
  for (int pp = 0; pp < S::NPASS; ++pp)
    for (int tb = 0; tb < S::TRB; ++tb)
      for (int s = 0; s < S::KBS; ++s)
        if (bmie_needed<C>(16 * S::TRB * pp + 16 * tb, s) || (pp == 0 && tb == 0)) m.rhs[pp][tb] |= 1ull << s;
  return m;
}

// LDS of one element's elimination.  Strides are padded so that the lanes of
// a wavefront hit distinct LDS banks.  T is the element type of the
// Gauss-Jordan elimination: double, or float for the FP32 + FP64 refinement
// form, which also eliminates identity tiles to obtain D'^-1.
template <class C, typename T = double>
struct SchurLds {
  using S = MfmaShape<C>;
  static constexpr int R = S::R, NPV = C::NPV, NPF = C::NPF, NC = C::NC;
  static constexpr bool UQ = C::UQ;
  static constexpr int RP = R + 1;                        // sPan / sZ: column t at t*RP
  static constexpr bool IR = sizeof(T) == 4;
  // tiles per elimination: D', the pass's RHS and, in FP32, the identity
  static constexpr int CTPX = S::CTP + (IR ? S::RT : 0);
  static constexpr int RSTR = 16 * CTPX + 1;              // sRows: pivot row t at t*RSTR
  static constexpr int NCOLP = C::NCOL | 1;               // sXt: RHS column cc at cc*NCOLP
  // more than one RHS pass: the panel transforms (Z, pivot rows) of the first
  // pass are kept and replayed on the later passes' RHS tiles, so D' is
  // factored once (u-only: with q the LDS would not leave room for one
  // wavefront per SIMD)
  static constexpr bool REPLAY = S::NPASS > 1 && !UQ && !IR;
  static constexpr bool RECORD = REPLAY;                  // Z of every panel kept
  static constexpr bool ROT = RECORD && (R & (R - 1)) == 0;
  static constexpr int ZSTR = ROT ? R : RP;
  // column t of the Z of panel pk, row `row`: recorded panels rotate the rows
  // by t when R is a power of two (bank spread without padding), else pad
  static HDGX_DEV constexpr int zidx(int pk, int t, int row) {
    return RECORD ? (pk * 4 + t) * ZSTR + (ROT ? ((row + t) & (R - 1)) : row) : t * RP + row;
  }
  // scratch inside sXt: the prologue's Cholesky factor of M; the elimination's
  // panel, pivot rows and (not recorded) Z, in T; in FP32 the staging of
  // D'^-1 (column j at j*NUP)
  static constexpr int NUP = C::NU | 1;
  static constexpr int NXT = NCOLP * 16;
  static constexpr int NSCR = (int)(((4 * RP + 4 * RSTR + (RECORD ? 0 : 4 * RP)) * sizeof(T) + 7) / 8);
  static constexpr int NINV = IR ? (C::NU * NUP * 4 + 7) / 8 : 0;
  static constexpr int NSX = NXT > NSCR ? (NXT > NINV ? NXT : NINV) : (NSCR > NINV ? NSCR : NINV);
  static_assert(!UQ || NPV * NPV <= NXT, "the factor of M fits the solution tile buffer");
  static constexpr int NNMAX = (NPV > 2 * NPF) ? NPV : 2 * NPF;
  static constexpr int KSMAX = (NNMAX * NC + 3) / 4;       // k-steps of the AE product, upper bound
  // sXt (the solution columns of one RHS tile) is live only in the output
  // phase; the scratch above shares its storage
  double sXt[NSX];
  T sZall[RECORD ? S::NPAN * 4 * ZSTR : 1];
  double sMiRq[UQ ? NPV * C::NCU * C::ND : 1];
  double sMiC[UQ ? NPV * NPV * C::ND : 1];
  double sMiE[UQ ? NPV * C::NDF * C::ND : 1];
  int sStep[R];
  int sPiv[RECORD ? S::NPAN * 4 : 1];
  int sAeNodes[S::RTH * NNMAX];
  int sAeNN[S::RTH];
  unsigned char sAeK[S::RTH * 4 * KSMAX];                // AE_NONE past the node set
};

// refinement residual of the FP32 form: A fragments of BD for k-steps
// s0 .. s0+BR-1 (row l%16 of each row tile, column 4s + l/16; zero past NCOL)
template <class C, int BR>
HDGX_DEV void ir_bd_frags(const double* __restrict__ BDe, int s0, int g, int c,
                          double (&a)[BR][MfmaShape<C>::RT]) {
  constexpr int NU = C::NU, NCOL = C::NCOL, RT = MfmaShape<C>::RT;
#pragma unroll
  for (int sb = 0; sb < BR; ++sb) {
    const int k = 4 * (s0 + sb) + g;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
      const int row = 16 * rt + c;
      a[sb][rt] = (k < NCOL && row < NU) ? BDe[row + NU * k] : 0.0;
    }
  }
}

// r -= BD(:, k-steps s0..s0+BR-1) X(those rows, :), X staged in LDS
template <class C, int BR>
HDGX_DEV void ir_residual_batch(const double* sXt, int s0, int g, int c, const double (&a)[BR][MfmaShape<C>::RT],
                                d4 (&r)[MfmaShape<C>::RT]) {
  constexpr int NCOL = C::NCOL, NCOLP = C::NCOL | 1, RT = MfmaShape<C>::RT;
  constexpr int KS = (NCOL + 3) / 4;
#pragma unroll
  for (int sb = 0; sb < BR; ++sb) {
    if (s0 + sb >= KS) break;
    const int k = 4 * (s0 + sb) + g;
    const double b = (k < NCOL) ? -sXt[k + NCOLP * c] : 0.0;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) r[rt] = mfma64(a[sb][rt], b, r[rt]);
  }
}

// AE = H + GK X: the column of GK and X that k-step s of AE row tile rta
// reads in lane group g (-1 past the tile's node set)
template <class C>
HDGX_DEV int ae_col(const unsigned char* sAeK, int rta, int s, int g) {
  constexpr int KSMAX = SchurLds<C>::KSMAX;
  if (s >= KSMAX) return -1;
  const int k = sAeK[rta * 4 * KSMAX + 4 * s + g];
  return k == AE_NONE ? -1 : k;
}

// Element e of the batch A (one 64-lane wavefront), LDS in L.
template <class C, typename T = double>
HDGX_DEV void schur_mfma_body(const SchurArgs& A, const int32_t* __restrict__ perm, const int64_t e,
                              SchurLds<C, T>& L) {
  using S = MfmaShape<C>;
  using LD = SchurLds<C, T>;
  using MT = Mfma<T>;
  typedef typename MT::V V;
  typedef typename MT::K K;
  constexpr int NU = C::NU, NH = C::NH, NCOL = C::NCOL, NPV = C::NPV, NCU = C::NCU, ND = C::ND;
  constexpr int NDF = C::NDF, NPF = C::NPF, NFE = C::NFE, NC = C::NC;
  constexpr bool UQ = C::UQ;
  constexpr int NRHS = S::NRHS, RT = S::RT, R = S::R, TRB = S::TRB, CTP = S::CTP, RTH = S::RTH;
  constexpr int NQ = UQ ? NPV * NCU * ND : 0;
  constexpr int NRQ = NCU * ND, NCC = NPV * ND, NEE = NDF * ND;
  static_assert(R <= 64, "one panel row per lane");
  static_assert(NPV * NC < AE_NONE, "AE column index fits a byte");

  const int lane = threadIdx.x;
  const int g = lane >> 4, c = lane & 15;
  if (e >= A.ne) return;

  constexpr int RP = LD::RP, RSTR = LD::RSTR, NCOLP = LD::NCOLP;
  constexpr bool IR = LD::IR, RECORD = LD::RECORD, REPLAY = LD::REPLAY;
  constexpr int CTPX = LD::CTPX, NUP = LD::NUP;
  constexpr int NNMAX = LD::NNMAX, KSMAX = LD::KSMAX;
  double* const sXt = L.sXt;
  T* const sPan = reinterpret_cast<T*>(L.sXt);
  T* const sRows = sPan + 4 * RP;
  T* const sZall = RECORD ? L.sZall : sRows + 4 * RSTR;
  int* const sStep = L.sStep;
  double* const sMf = L.sXt;
  double* const sMiRq = L.sMiRq;
  double* const sMiC = L.sMiC;
  double* const sMiE = L.sMiE;
  int* const sPiv = L.sPiv;
  int* const sAeNodes = L.sAeNodes;
  int* const sAeNN = L.sAeNN;
  unsigned char* const sAeK = L.sAeK;

  int st = 0;
  const double* BDe = A.BD + e * (int64_t)NU * NCOL;
  const double* Fe = A.F + e * (int64_t)NU * NH;
  const double* GKe = A.GK + e * (int64_t)NH * NCOL;
  const double* He = A.H + e * (int64_t)NH * NH;
  const int LDO = NPV * A.nco;   // DUDG / DUDG_DUH rows written (the first LDO of NCOL)
  double* Xe = A.dudg_duh + e * (int64_t)LDO * NH;

  // ---- node sets of the AE row tiles: the volume nodes of the faces met by
  // rows 16*rta .. 16*rta+15 (GK is face sparse); without perm (GK given by a
  // caller, possibly dense) every node.
  if (lane < RTH) {
    const int rta = lane;
    int nn = 0;
    if (perm) {
      const int j0 = (16 * rta) / (NCU * NPF);
      int j1 = (16 * rta + 15) / (NCU * NPF);
      if (j1 > NFE - 1) j1 = NFE - 1;
      for (int j = j0; j <= j1; ++j)
        for (int a = 0; a < NPF; ++a) {
          const int v = perm[a + NPF * j];
          bool seen = false;
          for (int q = 0; q < nn; ++q) seen |= (sAeNodes[rta * NNMAX + q] == v);
          if (!seen) sAeNodes[rta * NNMAX + nn++] = v;
        }
    } else {
      for (int v = 0; v < NPV; ++v) sAeNodes[rta * NNMAX + nn++] = v;
    }
    sAeNN[rta] = nn;
    for (int idx = 0; idx < 4 * KSMAX; ++idx)
      sAeK[rta * 4 * KSMAX + idx] =
          (unsigned char)((idx < nn * NC) ? sAeNodes[rta * NNMAX + idx % nn] + NPV * (idx / nn) : AE_NONE);
  }

  // ---- u,q prologue: M = U'U (dpotrf 'U'), M^-1 [Rq C E] (dpotrs)
  if constexpr (UQ) {
    for (int i = lane; i < NPV * NPV; i += 64) sMf[i] = A.M[e * NPV * NPV + i];
    __syncthreads();
    for (int j = 0; j < NPV; ++j) {
      if (lane == 0) {
        double v = sMf[j + NPV * j];
        if (!(v > 0.0)) st |= 1;
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
  }
  __syncthreads();

  // FP32 form: A fragments of D'^-1 (row 16 rt + l%16, column 4 s + l/16),
  // from the first pass's elimination
  constexpr int KSU = (NU + 3) / 4;
  T dinv[IR ? RT : 1][IR ? KSU : 1];
  for (int pass = 0; pass < S::NPASS; ++pass) {
    const int qbase = 16 * TRB * pass;       // first RHS column of this pass
    // FP64: every pass eliminates D' (later passes replay the recorded panels
    // when REPLAY); FP32: the first pass eliminates D', its RHS tiles and the
    // identity, the later passes' columns start from zero and are solved by
    // the refinement sweeps
    const bool full = (pass == 0) || !(RECORD || IR);
    const bool build = !IR || pass == 0;
    V acc[RT][CTPX];
    // ---- D' tiles: D, identity on the padded diagonal
    if (full)
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int ct = 0; ct < RT; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = 16 * rt + MT::row(i, g), col = 16 * ct + c;
          double v = 0.0;
          if (row < NU && col < NU) v = BDe[row + NU * col];
          else if (row == col) v = 1.0;
          acc[rt][ct][i] = (T)v;
        }
    // ---- RHS tiles: rhs0 | -F(:, trace column); FP32: identity tiles
    if constexpr (IR) {
      if (pass == 0)
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int ct = 0; ct < RT; ++ct)
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[rt][CTP + ct][i] = (16 * rt + MT::row(i, g) == 16 * ct + c) ? (T)1 : (T)0;
    }
    if (build)
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int tb = 0; tb < TRB; ++tb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = 16 * rt + MT::row(i, g), q = qbase + 16 * tb + c;
          double v = 0.0;
          if (row < NU && q < NRHS) v = (q == 0) ? A.Ru[e * NU + row] : -Fe[row + NU * rhs_t<C>(q)];
          acc[rt][RT + tb][i] = (T)v;
        }
    if (UQ && build) {
      // D' += B M^-1 C and RHS += B M^-1 E over k-steps of B's columns, A
      // fragments (columns of B) fetched a batch of k-steps ahead
      constexpr int BB = RT >= 4 ? 2 : 4;
      constexpr SkipMasks<C> SM = skip_masks<C>();
      // fully unrolled up to three row tiles; with four the unrolled loop
      // spills, so it stays rolled and tests the skip masks at run time
      constexpr int UNR = RT >= 4 ? 1 : (S::KBS + BB - 1) / BB;
#pragma unroll UNR
      for (int s0 = 0; s0 < S::KBS; s0 += BB) {
        T a[BB][RT];
#pragma unroll
        for (int sb = 0; sb < BB; ++sb)
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) {
            const int row = 16 * rt + c, ka = 4 * (s0 + sb) + g;     // A fragment: row l%16, k l/16
            a[sb][rt] = (s0 + sb < S::KBS && row < NU && ka < S::KB) ? (T)BDe[row + NU * (NU + ka)] : (T)0;
          }
#pragma unroll
        for (int sb = 0; sb < BB; ++sb) {
          const int s = s0 + sb;
          if (s >= S::KBS) break;
          const int kk = 4 * s + g;                                  // B-operand row of this lane
          const int am = kk % NPV, mm = (kk / NPV) % NCU, dm = kk / (NPV * NCU);
#pragma unroll
          for (int ct = 0; ct < RT; ++ct) {
            if (!full || !((SM.ct[ct] >> s) & 1ull)) continue;
            const int col = 16 * ct + c;
            T b = 0;
            if (kk < S::KB && col < NU && col / NPV == mm) b = (T)sMiC[am + NPV * (col % NPV) + NPV * NPV * dm];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) acc[rt][ct] = MT::op(a[sb][rt], b, acc[rt][ct]);
          }
#pragma unroll
          for (int tb = 0; tb < TRB; ++tb) {
            if (!((SM.rhs[pass][tb] >> s) & 1ull)) continue;
            const int q = qbase + 16 * tb + c;
            T b = 0;
            // column 0: rhs0 = Ru - B M^-1 Rq; columns q >= 1: B M^-1 E - F
            if (kk < S::KB) {
              if (q == 0) b = (T)-sMiRq[kk];
              else if (q < NRHS && rhs_k<C>(q) == mm) b = (T)sMiE[am + NPV * rhs_jf<C>(q) + NPV * NDF * dm];
            }
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) acc[rt][RT + tb] = MT::op(a[sb][rt], b, acc[rt][RT + tb]);
          }
        }
      }
    }

    // ---- Gauss-Jordan with partial pivoting, panels of four columns
    if (full) {
    if (lane < R) sStep[lane] = -1;
    bool done = false;                        // lane r < R: row r already a pivot row
    __syncthreads();
    // panels: unrolled over the panel's column tile (register indices are
    // then static), a runtime loop over the four panels inside the tile
#pragma unroll
    for (int ct0 = 0; ct0 < RT; ++ct0)
#pragma unroll 1
    for (int sub = 0; sub < 4; ++sub) {
      const int c0 = 4 * sub, k0 = 16 * ct0 + c0;
      if (k0 >= NU) break;
      // panel columns -> LDS, one row per lane
      {
        const int t = c - c0;
        if (t >= 0 && t < 4)
#pragma unroll
          for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int i = 0; i < 4; ++i) sPan[t * RP + 16 * rt + MT::row(i, g)] = acc[rt][ct0][i];
      }
      __syncthreads();
      T pr[4], z[4] = {0, 0, 0, 0};
#pragma unroll
      for (int t = 0; t < 4; ++t) pr[t] = (lane < R) ? sPan[t * RP + lane] : (T)0;
      int pvt[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        // |pr| as an ordered key whose low 6 bits carry 63-lane, so one
        // max-reduction finds the largest magnitude and, on ties, the lowest lane
        K key = 0;
        if (lane < R && !done) key = (MT::key(fabs(pr[t])) & ~(K)63) | (K)(63 - lane);
        key = wave_key_max(key);
        if ((key >> 6) == 0) st |= 2;
        const int p = __builtin_amdgcn_readfirstlane(63 - (int)(key & 63));
        pvt[t] = p;
        const T pv = lane_bcast(pr[t], p);
        const T ip = (T)1 / pv;
        z[t] = (lane == p) ? (T)1 : (T)0;
        T ppr[4], pz[4];
#pragma unroll
        for (int t2 = 0; t2 < 4; ++t2) { ppr[t2] = lane_bcast(pr[t2], p); pz[t2] = lane_bcast(z[t2], p); }
        if (lane == p) {
#pragma unroll
          for (int t2 = 0; t2 < 4; ++t2) { pr[t2] *= ip; z[t2] *= ip; }
          done = true;
        } else {
          const T m = pr[t] * ip;
#pragma unroll
          for (int t2 = 0; t2 < 4; ++t2) { pr[t2] -= m * ppr[t2]; z[t2] -= m * pz[t2]; }
        }
        if (lane == 0) sStep[p] = k0 + t;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) if (lane == pvt[t]) z[t] -= (T)1;
      const int pk = k0 / 4;
      if (lane < R) {
#pragma unroll
        for (int t = 0; t < 4; ++t) sZall[LD::zidx(pk, t, lane)] = z[t];
      }
      if (RECORD && lane < 4)
        sPiv[pk * 4 + lane] = pvt[0] * (lane == 0) + pvt[1] * (lane == 1) + pvt[2] * (lane == 2) + pvt[3] * (lane == 3);
      // original (pre-panel) pivot rows, columns from tile ct0 on: the tile
      // row and register holding pivot row p are uniform, so the selection is
      // a scalar branch; lane group grp(p) owns the row
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int p = pvt[t];
        const int sel = 4 * (p >> 4) + MT::reg(p & 15);     // 4*rt + i
        T tmp[CTPX];
#pragma unroll
        for (int ct = 0; ct < CTPX; ++ct) tmp[ct] = 0;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (sel == 4 * rt + i)
#pragma unroll
              for (int ct = 0; ct < CTPX; ++ct)
                if (ct >= ct0) tmp[ct] = acc[rt][ct][i];
        if (g == MT::grp(p & 15))
#pragma unroll
          for (int ct = 0; ct < CTPX; ++ct)
            if (ct >= ct0) sRows[t * RSTR + 16 * ct + c] = tmp[ct];
      }
      __syncthreads();
      // S += Z * S[pivot rows, :]
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        const T a = sZall[LD::zidx(pk, g, 16 * rt + c)];
#pragma unroll
        for (int ct = 0; ct < CTPX; ++ct)
          if (ct >= ct0) acc[rt][ct] = MT::op(a, sRows[g * RSTR + 16 * ct + c], acc[rt][ct]);
      }
      __syncthreads();
    }
    if constexpr (IR) {
      // the identity tiles now hold D'^-1 with row k at pivot row p_k:
      // staged un-permuted through LDS into the A fragments
      T* const sInv = reinterpret_cast<T*>(sXt);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < RT; ++ct)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int k = sStep[16 * rt + MT::row(i, g)], col = 16 * ct + c;
            if (k >= 0 && k < NU && col < NU) sInv[k + NUP * col] = acc[rt][CTP + ct][i];
          }
      __syncthreads();
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int s = 0; s < KSU; ++s) {
          const int row = 16 * rt + c, kk = 4 * s + g;
          dinv[rt][s] = (row < NU && kk < NU) ? sInv[row + NUP * kk] : (T)0;
        }
      __syncthreads();
    }
    } else if (REPLAY) {
      // replay the recorded panel transforms on this pass's RHS tiles
#pragma unroll 1
      for (int pk = 0; pk < S::NPAN; ++pk) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int p = __builtin_amdgcn_readfirstlane(sPiv[pk * 4 + t]);
          const int sel = 4 * (p >> 4) + MT::reg(p & 15);
          T tmp[TRB];
#pragma unroll
          for (int tb = 0; tb < TRB; ++tb) tmp[tb] = 0;
#pragma unroll
          for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int i = 0; i < 4; ++i)
              if (sel == 4 * rt + i)
#pragma unroll
                for (int tb = 0; tb < TRB; ++tb) tmp[tb] = acc[rt][RT + tb][i];
          if (g == MT::grp(p & 15))
#pragma unroll
            for (int tb = 0; tb < TRB; ++tb) sRows[t * RSTR + 16 * (RT + tb) + c] = tmp[tb];
        }
        __syncthreads();
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
          const T a = sZall[LD::zidx(pk, g, 16 * rt + c)];
#pragma unroll
          for (int tb = 0; tb < TRB; ++tb)
            acc[rt][RT + tb] = MT::op(a, sRows[g * RSTR + 16 * (RT + tb) + c], acc[rt][RT + tb]);
        }
        __syncthreads();
      }
    }

    // ---- per RHS tile: u rows by pivot order, q rows, outputs, AE/FE columns
#pragma unroll
    for (int tb = 0; tb < TRB; ++tb) {
      const int q0 = qbase + 16 * tb;
      if (q0 >= NRHS) break;
      double xmx = 0.0;
      if (build) {
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int k = sStep[16 * rt + MT::row(i, g)];
            if (k >= 0 && k < NU) {
              const double v = (double)acc[rt][RT + tb][i];
              sXt[k + NCOLP * c] = v;
              xmx = fmax(xmx, fabs(v));
            }
          }
      } else {
        for (int r = g; r < NU; r += 4) sXt[r + NCOLP * c] = 0.0;
      }
      __syncthreads();
      // q rows; in FP32 then FP64 refinement sweeps.  With the q rows of X
      // consistent with its u rows, the residual of the reduced system
      // D' x = rhs is that of the u equations of the element system,
      // r = [Ru | -F] - BD X, formed here in FP64 from BD and the FP64
      // solution; the correction is the FP32 product D'^-1 r.  The sweeps
      // stop once the predicted error |d|^2 / |d_prev| (the last contraction
      // applied once more) is below 1e-13 |x|; columns starting from zero
      // spend their first sweep on the solve itself.
      double dprev = IR ? wave_max(xmx) : 0.0;
      bool conv = false;
#pragma unroll 1
      for (int sweep = 0;; ++sweep) {
        if constexpr (UQ) {
          // dq = MiRq + MiC du ; dlq = MiC dlu - MiE (x) I   (schur_uq.c:146-186):
          // the q rows n + NPV (m + NCU dd) of component m and direction dd
          // per lane group, with that component's du of the lane's column
          // loaded once
          const int q = q0 + c;
          for (int bq = g; bq < NCU * ND; bq += 4) {
            const int m = bq % NCU, dd = bq / NCU;
            double xu[NPV];
#pragma unroll
            for (int a = 0; a < NPV; ++a) xu[a] = sXt[a + NPV * m + NCOLP * c];
            const bool diag = q >= 1 && q < NRHS && rhs_k<C>(q) == m;
            const int jf = diag ? rhs_jf<C>(q) : 0;
#pragma unroll
            for (int n = 0; n < NPV; ++n) {
              double v = 0.0;
#pragma unroll
              for (int a = 0; a < NPV; ++a) v += sMiC[n + NPV * a + NPV * NPV * dd] * xu[a];
              if (q == 0) v += sMiRq[n + NPV * m + NPV * NCU * dd];
              else if (diag) v -= sMiE[n + NPV * jf + NPV * NDF * dd];
              sXt[NU + n + NPV * (m + NCU * dd) + NCOLP * c] = v;
            }
          }
          __syncthreads();
        }
        if constexpr (!IR) {
          break;
        } else {
          if (conv) break;
          if (sweep == IR_SWEEPS + (build ? 0 : 1)) { st |= 8; break; }
          const int q = q0 + c;
          d4 r[RT];
#pragma unroll
          for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int row = 16 * rt + 4 * i + g;
              double v = 0.0;
              if (row < NU && q < NRHS) v = (q == 0) ? A.Ru[e * NU + row] : -Fe[row + NU * rhs_t<C>(q)];
              r[rt][i] = v;
            }
          // k-steps over the columns of BD in batches of BR, the A fragments
          // of the next batch loaded while this one multiplies
          constexpr int KS = (NCOL + 3) / 4, BR = 4;
          double a0[BR][RT], a1[BR][RT];
          ir_bd_frags<C, BR>(BDe, 0, g, c, a0);
#pragma unroll 1
          for (int s0 = 0; s0 < KS; s0 += 2 * BR) {
            ir_bd_frags<C, BR>(BDe, s0 + BR, g, c, a1);
            ir_residual_batch<C, BR>(sXt, s0, g, c, a0, r);
            ir_bd_frags<C, BR>(BDe, s0 + 2 * BR, g, c, a0);
            ir_residual_batch<C, BR>(sXt, s0 + BR, g, c, a1, r);
          }
          // d = D'^-1 r in FP32: r in the FP64 accumulator layout is the B
          // operand directly (k-step s = rows 4s + l/16 = register s%4 of
          // row tile s/4)
          V dl[RT];
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) dl[rt] = V{0, 0, 0, 0};
#pragma unroll
          for (int s = 0; s < KSU; ++s) {
            const T b = (T)r[s / 4][s % 4];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) dl[rt] = MT::op(dinv[rt][s], b, dl[rt]);
          }
          double dmx = 0.0, xm = 0.0;
#pragma unroll
          for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int row = 16 * rt + MT::row(i, g);
              if (row < NU) {
                const double d = (double)dl[rt][i];
                const double v = sXt[row + NCOLP * c] + d;
                sXt[row + NCOLP * c] = v;
                dmx = fmax(dmx, fabs(d));
                xm = fmax(xm, fabs(v));
              }
            }
          __syncthreads();
          dmx = wave_max(dmx);
          xm = wave_max(xm);
          conv = dmx == 0.0 || dmx * dmx <= 1e-13 * xm * dprev;
          dprev = dmx;
        }
      }
      // DUDG / DUDG_DUH columns of this tile (coalesced over rows); the
      // rows kept are all NCOL (the usual case, a compile-time bound) or the
      // first LDO
      auto put_cols = [&](const int ldo, const int nrow) {
        for (int cc = 0; cc < 16; ++cc) {
          const int q = q0 + cc;
          if (q >= NRHS) break;
          for (int r = lane; r < nrow; r += 64) {
            const double v = sXt[r + NCOLP * cc];
            if (!isfinite(v)) st |= 4;
            if (q == 0) A.dudg[e * ldo + r] = v;
            else Xe[r + (int64_t)ldo * rhs_t<C>(q)] = v;
          }
        }
      };
      if (LDO == NCOL) put_cols(NCOL, NCOL);
      else put_cols(LDO, LDO);
      // AE(:, tile) = H + GK X, FE = Rh - GK x (schur_uq.c:279-285); all AE row
      // tiles advance together, fragments fetched a batch of k-steps ahead
      {
        const int q = q0 + c;
        d4 ae[RTH];
#pragma unroll
        for (int rta = 0; rta < RTH; ++rta)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = 16 * rta + 4 * i + g;
            ae[rta][i] = (row < NH && q >= 1 && q < NRHS) ? He[row + NH * rhs_t<C>(q)] : 0.0;
          }
        constexpr int BATCH = RTH >= 4 ? 2 : 4;
        int klen[RTH], kend = 0;     // k-steps that meet a tile's node set
#pragma unroll
        for (int rta = 0; rta < RTH; ++rta) {
          klen[rta] = __builtin_amdgcn_readfirstlane(sAeNN[rta] * NC);
          kend = max(kend, (klen[rta] + 3) / 4);
        }
#pragma unroll 1
        for (int s0 = 0; s0 < kend; s0 += BATCH) {
          double a[BATCH][RTH], b[BATCH][RTH];
#pragma unroll
          for (int sb = 0; sb < BATCH; ++sb)
#pragma unroll
            for (int rta = 0; rta < RTH; ++rta) {
              const int k = ae_col<C>(sAeK, rta, s0 + sb, g);
              const int row_a = 16 * rta + c;
              a[sb][rta] = (k >= 0 && row_a < NH) ? GKe[row_a + NH * k] : 0.0;
              b[sb][rta] = k >= 0 ? sXt[k + NCOLP * c] : 0.0;
            }
#pragma unroll
          for (int sb = 0; sb < BATCH; ++sb)
#pragma unroll
            for (int rta = 0; rta < RTH; ++rta)
              if (s0 + sb < KSMAX && 4 * (s0 + sb) < klen[rta]) ae[rta] = mfma64(a[sb][rta], b[sb][rta], ae[rta]);
        }
#pragma unroll
        for (int rta = 0; rta < RTH; ++rta)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = 16 * rta + 4 * i + g;
            if (row < NH && q < NRHS) {
              double v;
              if (q == 0) { v = A.Rh[e * NH + row] - ae[rta][i]; A.fe[e * NH + row] = v; }
              else { v = ae[rta][i]; A.ae[e * (int64_t)NH * NH + row + NH * rhs_t<C>(q)] = v; }
              if (!isfinite(v)) st |= 4;
            }
          }
      }
      __syncthreads();
    }
  }
  st = wave_or(st);
  if (lane == 0 && A.status) A.status[e] = st;
}

// two wavefronts per SIMD where the system fits: the compiler then keeps the
// accumulator tiles in VGPRs within 256 registers (15% faster on NS p=3 than
// one wavefront per SIMD with AGPR accumulators)
template <class C>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MfmaShape<C>::WPE, MfmaShape<C>::WPE))) void schur_mfma_kernel(SchurArgs A, const int32_t* __restrict__ perm) {
  __shared__ SchurLds<C> lds;
  schur_mfma_body<C>(A, perm, blockIdx.x, lds);
}

// FP32 MFMA Gauss-Jordan with FP64 refinement of every solution column
template <class C>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MfmaShape<C>::WPE_IR, MfmaShape<C>::WPE_IR))) void schur_mfma_ir_kernel(SchurArgs A, const int32_t* __restrict__ perm) {
  __shared__ SchurLds<C, float> lds;
  schur_mfma_body<C, float>(A, perm, blockIdx.x, lds);
}

}  // namespace hdgx
