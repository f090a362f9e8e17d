// FP64 element elimination for the systems of one to four 16-row tiles
// (kernel/schur_u.c:42-163, kernel/schur_uq.c:40-287): every 2D order p <= 3,
// tetrahedra p = 1 and, u-only at four row tiles, 2D Euler p = 4 and Euler
// tetrahedra p = 2.  One wavefront per element, FP64 MFMA throughout.
//
// The element system lives in v_mfma_f64_16x16x4f64 accumulator tiles (lane
// l = 16g + c, register i of a tile: row 4i + g, column c) as
// [D' | du | trace RHS]:
//   * D' = D (+ B M^-1 C) fills columns 0..NU-1 of the first RT tiles; the
//     padding column NU of those tiles carries du's right-hand side
//     Ru (- B M^-1 Rq), so du costs no tile of its own (NS p = 3: three RHS
//     tiles instead of four);
//   * the trace RHS -F (+ B M^-1 E), trace column j in tile j / 16, in the
//     order j = jf + NDF * k (face node jf, component k), so a tile meets few
//     components.
// Gauss-Jordan with partial pivoting in panels of four columns (the panel
// factored one row per lane, a DPP argmax for each pivot, the rank-4 update
// S <- S + Z S[pivot rows, :] applied by MFMA), then per RHS tile: the q rows
// M^-1 (C x) - M^-1 E as MFMA products, DUDG_DUH, and AE = H + GK X by MFMA
// over the face-sparse columns of GK (faceint.m:354-358) whose A fragments
// are loaded once into the registers D' frees, written through LDS in whole
// columns.  du's q rows, DUDG and FE = Rh - GK x run once, on the VALU.
//
// AE k-step tables come precomputed with the master (hdgx_api.hip,
// ae_table); without them (hdgx_schur_u / hdgx_schur_uq, GK given by a
// caller) every column of GK and H is read.
#include <hip/hip_runtime.h>

#include "hdgx_device.h"
#include "hdgx_kernels.h"
#include "hdgx_shapes.h"

namespace hdgx {

// does k-step s of B (columns 4s..4s+3, column K = a + NPV*m + NPV*NCU*d)
// meet a nonzero block of M^-1 C in D' column tile ct, or the du column?
template <class C>
HDGX_DEV constexpr bool s64_bmic(int ct, int s) {
  using S = S64Shape<C>;
  if (ct == S::CTD) return true;                 // -B M^-1 Rq is dense
  for (int kk = 4 * s; kk < 4 * s + 4 && kk < S::KB; ++kk) {
    const int m = (kk / C::NPV) % C::NCU;
    for (int col = 16 * ct; col < 16 * ct + 16 && col < C::NU; ++col)
      if (col / C::NPV == m) return true;
  }
  return false;
}

// the same for RHS tile tb: trace columns j of component j / NDF
template <class C>
HDGX_DEV constexpr bool s64_bmie(int tb, int s) {
  using S = S64Shape<C>;
  for (int kk = 4 * s; kk < 4 * s + 4 && kk < S::KB; ++kk) {
    const int m = (kk / C::NPV) % C::NCU;
    for (int j = 16 * tb; j < 16 * tb + 16 && j < C::NH; ++j)
      if (j / C::NDF == m) return true;
  }
  return false;
}

template <class C>
struct S64Masks {
  uint64_t ct[S64Shape<C>::RT];
  uint64_t rhs[S64Shape<C>::TRH];
};

template <class C>
HDGX_DEV constexpr S64Masks<C> s64_masks() {
  using S = S64Shape<C>;
  static_assert(S::KBS <= 64, "k-steps of B fit a 64-bit mask");
  S64Masks<C> m{};
  for (int ct = 0; ct < S::RT; ++ct)
    for (int s = 0; s < S::KBS; ++s)
      if (s64_bmic<C>(ct, s)) m.ct[ct] |= 1ull << s;
  for (int tb = 0; tb < S::TRH; ++tb)
    for (int s = 0; s < S::KBS; ++s)
      if (s64_bmie<C>(tb, s)) m.rhs[tb] |= 1ull << s;
  return m;
}

// trace column (k + NCU jf) of RHS column j = jf + NDF k
template <class C> HDGX_DEV constexpr int s64_t(int j) { return j / C::NDF + C::NCU * (j % C::NDF); }

template <class C>
struct S64Lds {
  using S = S64Shape<C>;
  static constexpr int R = S::R, RP = R + 1, CT = S::CTW, NPV = C::NPV;
  static constexpr int RSTR = 16 * CT + 1;       // pivot row t at t*RSTR
  static constexpr int NCOLP = C::NCOL | 1;      // solution column cc at cc*NCOLP
  static constexpr int NHP = C::NH | 1;          // AE staging column cc at cc*NHP
  static constexpr int NXT = NCOLP * 16;
  static constexpr int NSCR = 8 * RP + 4 * RSTR; // panel, Z, pivot rows
  static constexpr int M1 = NXT > NSCR ? NXT : NSCR;
  static constexpr int M2 = NHP * 16 > NPV * NPV ? NHP * 16 : NPV * NPV;
  static constexpr int NSX = M1 > M2 ? M1 : M2;
  static constexpr bool UQ = C::UQ;
  static constexpr int NTAB = S::TRH * (1 + 4 * S::KSMAX);
  // solutions of one RHS tile; the scratch of the other phases shares it
  double sX[NSX];
  double sMiRq[UQ ? NPV * C::NCU * C::ND : 1];
  double sMiC[UQ ? NPV * NPV * C::ND : 1];
  double sMiE[UQ ? NPV * C::NDF * C::ND : 1];
  int sStep[R];
  // klen[TRH], the k-step table, then (tile layout) the node-on-face table
  // loc(v, f) of F's face blocks
  unsigned char sTab[(NTAB + C::NPV * C::NFE + 7) / 8 * 8];
};

// Element e of A (one 64-lane wavefront), LDS in L.  TILE: F and H are face
// blocks and GK is in fragment order (WsLayout, hdgx_shapes.h), and A.aek is
// the master's AE table followed by its node-on-face table; otherwise every
// array has the reference layout.  With four row tiles the RHS tiles are split
// over S::NW wavefronts: `part` takes RHS tiles part*TRW .. part*TRW + TRW - 1
// (AE / DUDG_DUH columns), part 0 also du, DUDG and FE; the status word is OR-ed into A.status[e], which the launch
// zeroes first.
template <class C, bool TILE = false>
HDGX_DEV void schur64_body(const SchurArgs& A, const int64_t e, const int part, S64Lds<C>& L) {
  using S = S64Shape<C>;
  using WL = WsLayout<C>;
  using LD = S64Lds<C>;
  static_assert(S::USE, "schur64: one to four row tiles (four: u only), NU below the padded size");
  constexpr int NU = S::NU, NH = S::NH, NCOL = S::NCOL, NPV = S::NPV, NCU = S::NCU, ND = S::ND;
  constexpr int NDF = S::NDF, NPF = S::NPF, NC = S::NC, RT = S::RT, R = S::R, TRH = S::TRH, CT = S::CTW;
  constexpr int NW = S::NW, TRW = S::TRW;       // RHS tiles split over NW wavefronts, TRW each
  static_assert(NW == 1 || !C::UQ, "the RHS split is for u-only systems");
  constexpr int DU = S::DU, CTD = S::CTD, KSMAX = S::KSMAX;
  constexpr bool UQ = C::UQ;
  constexpr int NQ = UQ ? NPV * NCU * ND : 0;
  constexpr int NRQ = NCU * ND, NCC = NPV * ND, NEE = NDF * ND;
  constexpr int RP = LD::RP, RSTR = LD::RSTR, NCOLP = LD::NCOLP, NHP = LD::NHP;
  constexpr int NB = NCU * NPF;                 // trace rows of one face
  static_assert(R <= 64, "one panel row per lane");

  const int lane = threadIdx.x;
  const int g = lane >> 4, c = lane & 15;
  if (e >= A.ne) return;
  const bool sparse = TILE || A.aek != nullptr;
  double* const sX = L.sX;
  int* const sStep = L.sStep;
  int st = 0;
  const double* BDe = A.BD + e * (int64_t)NU * NCOL;
  const double* Fe = A.F + e * (int64_t)WL::f(TILE);
  const double* GKe = A.GK + e * (int64_t)WL::gk(TILE);
  const double* He = A.H + e * (int64_t)WL::h(TILE);
  const int LDO = NPV * A.nco;                   // DUDG / DUDG_DUH rows kept

  // ---- D' and du's right-hand side in column DU, loaded first so that their
  // latency overlaps the prologue
  d4 acc[RT][CT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < RT; ++ct)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * rt + 4 * i + g, col = 16 * ct + c;
        double v = 0.0;
        if (row < NU) {
          if (col < NU) v = BDe[row + NU * col];
          else if (col == DU) v = A.Ru[e * NU + row];
        }
        acc[rt][ct][i] = v;
      }
  // F's face blocks (tile layout), read whole with coalesced loads and put in
  // sX past the u,q prologue's Cholesky factor, where the -F gather below
  // reads them
  constexpr int NFB = C::NFE * NB * NB, FOFF = UQ ? NPV * NPV + NPV : 0;
  constexpr bool FST = TILE;
  static_assert(!TILE || FOFF + NFB <= LD::NSX, "F's face blocks fit in sX");
  constexpr int IFS = FST ? (NFB + 63) / 64 : 1;
  double fr[IFS];
  if constexpr (FST) {
#pragma unroll
    for (int it = 0; it < IFS; ++it) fr[it] = lane + 64 * it < NFB ? Fe[lane + 64 * it] : 0.0;
  }

  // ---- AE k-step table: the face-sparse one of the master, or every column
  if (sparse) {
    for (int i = lane; i < LD::NTAB + (TILE ? NPV * C::NFE : 0); i += 64) L.sTab[i] = A.aek[i];
  } else {
    for (int i = lane; i < LD::NTAB; i += 64) {
      int v;
      if (i < TRH) v = NCOL;
      else {
        const int k = (i - TRH) % (4 * KSMAX);
        v = k < NCOL ? k : AE_NONE;
      }
      L.sTab[i] = (unsigned char)v;
    }
  }

  // ---- u,q prologue: M = U'U (dpotrf 'U'), M^-1 [Rq C E] (dpotrs)
  if constexpr (UQ) {
    double* const sMf = sX;
    for (int i = lane; i < NPV * NPV; i += 64) sMf[i] = A.M[e * NPV * NPV + i];
    __syncthreads();
    for (int j = 0; j < NPV; ++j) {
      if (lane == 0) {
        const double v = sMf[j + NPV * j];
        if (!(v > 0.0)) st |= 1;
        const double r = sqrt(v);
        sMf[j + NPV * j] = r;
        sMf[NPV * NPV + j] = 1.0 / r;           // the reciprocal diagonal for the solves
      }
      __syncthreads();
      if (lane > j && lane < NPV) sMf[j + NPV * lane] *= sMf[NPV * NPV + j];
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
      if (col < NRQ) { src = A.Rq + e * NPV * NRQ + NPV * col; dst = L.sMiRq + NPV * col; }
      else if (col < NRQ + NCC) { src = A.C + e * NPV * NCC + NPV * (col - NRQ); dst = L.sMiC + NPV * (col - NRQ); }
      else { src = A.E + e * NPV * NEE + NPV * (col - NRQ - NCC); dst = L.sMiE + NPV * (col - NRQ - NCC); }
      double x[NPV];
#pragma unroll
      for (int i = 0; i < NPV; ++i) x[i] = src[i];
#pragma unroll
      for (int i = 0; i < NPV; ++i) {
        double s = x[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= sMf[k + NPV * i] * x[k];
        x[i] = s * sMf[NPV * NPV + i];
      }
#pragma unroll
      for (int i = NPV - 1; i >= 0; --i) {
        double s = x[i];
#pragma unroll
        for (int k = i + 1; k < NPV; ++k) s -= sMf[i + NPV * k] * x[k];
        x[i] = s * sMf[NPV * NPV + i];
      }
#pragma unroll
      for (int i = 0; i < NPV; ++i) dst[i] = x[i];
    }
  }
  if constexpr (FST) {
#pragma unroll
    for (int it = 0; it < IFS; ++it)
      if (lane + 64 * it < NFB) sX[FOFF + lane + 64 * it] = fr[it];
  }
  __syncthreads();

  // the trace columns -F of this wavefront's RHS tiles
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int tb = 0; tb < TRW; ++tb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int tbg = part * TRW + tb;
        const int row = 16 * rt + 4 * i + g, j = 16 * tbg + c;
        if constexpr (TILE) {
          // F's face block f of trace column j = (f NPF + b) + NDF k: row
          // va + NPV m at q = loc(va, f) + NPF m, column k + NCU b
          const int jf = j % NDF, k = j / NDF, f = jf / NPF, b = jf % NPF, va = row % NPV, m = row / NPV;
          const int la = row < NU && j < NH ? L.sTab[LD::NTAB + va + NPV * f] : 255;
          const double x = sX[FOFF + (la != 255 ? f * NB * NB + la + NPF * m + NB * (k + NCU * b) : 0)];
          acc[rt][RT + tb][i] = la != 255 ? -x : 0.0;
        } else {
          acc[rt][RT + tb][i] = (tbg < TRH && row < NU && j < NH) ? -Fe[row + NU * s64_t<C>(j)] : 0.0;
        }
      }
  if constexpr (FST) __syncthreads();          // sX is the elimination's scratch below
  // ---- D' += B M^-1 C, du's column -= B M^-1 Rq, RHS += B M^-1 E, over
  // k-steps of B's columns whose blocks are nonzero (compile-time masks), A
  // fragments (columns of B) fetched a batch of k-steps ahead
  if constexpr (UQ) {
    constexpr int BB = 4;
    constexpr S64Masks<C> SM = s64_masks<C>();
    // B's A fragments a batch of k-steps ahead of the products that use them
    auto load_b = [&](int s0, double (&a)[BB][RT]) {
#pragma unroll
      for (int sb = 0; sb < BB; ++sb)
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
          const int row = 16 * rt + c, ka = 4 * (s0 + sb) + g;
          a[sb][rt] = (s0 + sb < S::KBS && row < NU && ka < S::KB) ? BDe[row + NU * (NU + ka)] : 0.0;
        }
    };
    double a[BB][RT], an[BB][RT];
    load_b(0, a);
#pragma unroll
    for (int s0 = 0; s0 < S::KBS; s0 += BB) {
      if (s0 + BB < S::KBS) load_b(s0 + BB, an);
#pragma unroll
      for (int sb = 0; sb < BB; ++sb) {
        const int s = s0 + sb;
        if (s >= S::KBS) break;
        const int kk = 4 * s + g;                                  // B-operand row of this lane
        const int am = kk % NPV, mm = (kk / NPV) % NCU, dm = kk / (NPV * NCU);
#pragma unroll
        for (int ct = 0; ct < RT; ++ct) {
          if (!((SM.ct[ct] >> s) & 1ull)) continue;
          const int col = 16 * ct + c;
          double b = 0.0;
          if (kk < S::KB) {
            if (col < NU && col / NPV == mm) b = L.sMiC[am + NPV * (col % NPV) + NPV * NPV * dm];
            else if (col == DU) b = -L.sMiRq[kk];
          }
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) acc[rt][ct] = mfma64(a[sb][rt], b, acc[rt][ct]);
        }
#pragma unroll
        for (int tb = 0; tb < TRH; ++tb) {
          if (!((SM.rhs[tb] >> s) & 1ull)) continue;
          const int j = 16 * tb + c;
          double b = 0.0;
          if (kk < S::KB && j < NH && j / NDF == mm) b = L.sMiE[am + NPV * (j % NDF) + NPV * NDF * dm];
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) acc[rt][RT + tb] = mfma64(a[sb][rt], b, acc[rt][RT + tb]);
        }
      }
#pragma unroll
      for (int sb = 0; sb < BB; ++sb)
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) a[sb][rt] = an[sb][rt];
    }
  }

  // ---- Gauss-Jordan with partial pivoting, panels of four columns
  {
    double* const sPan = sX;
    double* const sZ = sX + 4 * RP;
    double* const sRows = sX + 8 * RP;
    if (lane < R) sStep[lane] = -1;
    bool done = false;                        // lane r < R: row r already a pivot row
    __syncthreads();
#pragma unroll
    for (int ct0 = 0; ct0 < RT; ++ct0)
#pragma unroll 1
      for (int sub = 0; sub < 4; ++sub) {
        const int c0 = 4 * sub, k0 = 16 * ct0 + c0;
        if (k0 >= NU) break;
        const int np = NU - k0 < 4 ? NU - k0 : 4;   // pivots of this panel (fewer in a last, partial one)
        {
          const int t = c - c0;
          if (t >= 0 && t < 4)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
#pragma unroll
              for (int i = 0; i < 4; ++i) sPan[t * RP + 16 * rt + 4 * i + g] = acc[rt][ct0][i];
        }
        __syncthreads();
        double pr[4], z[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < 4; ++t) pr[t] = (lane < R) ? sPan[t * RP + lane] : 0.0;
        int pvt[4] = {-1, -1, -1, -1};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if (t >= np) break;
          // |pr| as an ordered 32-bit key whose low 6 bits carry 63 - lane:
          // one max-reduction finds the largest magnitude (piv_key), ties to
          // the lowest lane
          unsigned key = 0;
          if (lane < R && !done) key = piv_key(pr[t], lane);
          key = wave_key_max(key);
          if ((key >> 6) == 0) st |= 2;
          const int p = __builtin_amdgcn_readfirstlane(63 - (int)(key & 63));
          pvt[t] = p;
          const double pv = lane_bcast(pr[t], p);
          const double ip = recip(pv);
          z[t] = (lane == p) ? 1.0 : 0.0;
          // the pivot row's later panel columns and its multipliers of the
          // earlier pivots (its z(t) is 1, later z 0; a row's earlier columns
          // are not read again)
          double ppr[4], pz[4];
#pragma unroll
          for (int t2 = 0; t2 < 4; ++t2) {
            ppr[t2] = t2 > t ? lane_bcast(pr[t2], p) : 0.0;
            pz[t2] = t2 < t ? lane_bcast(z[t2], p) : (t2 == t ? 1.0 : 0.0);
          }
          if (lane == p) {
#pragma unroll
            for (int t2 = 0; t2 < 4; ++t2) { pr[t2] *= ip; z[t2] *= ip; }
            done = true;
          } else {
            const double m = pr[t] * ip;
#pragma unroll
            for (int t2 = 0; t2 < 4; ++t2) {
              if (t2 > t) pr[t2] -= m * ppr[t2];
              z[t2] -= m * pz[t2];
            }
          }
          if (lane == 0) sStep[p] = k0 + t;
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) if (t < np && lane == pvt[t]) z[t] -= 1.0;
        if (lane < R) {
#pragma unroll
          for (int t = 0; t < 4; ++t) sZ[t * RP + lane] = z[t];
        }
        // original (pre-panel) pivot rows, columns from tile ct0 on: lane group
        // p & 3 owns row p in register (p & 15) >> 2 of row tile p >> 4; the
        // rows of a partial panel's missing pivots are zero (their Z column is)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if (t >= np) {
            if (g == 0)
#pragma unroll
              for (int ct = 0; ct < CT; ++ct)
                if (ct >= ct0) sRows[t * RSTR + 16 * ct + c] = 0.0;
            continue;
          }
          const int p = pvt[t];
          const int sel = 4 * (p >> 4) + ((p & 15) >> 2);
          double tmp[CT];
#pragma unroll
          for (int ct = 0; ct < CT; ++ct) tmp[ct] = 0.0;
#pragma unroll
          for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int i = 0; i < 4; ++i)
              if (sel == 4 * rt + i)
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
                  if (ct >= ct0) tmp[ct] = acc[rt][ct][i];
          if (g == (p & 3))
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
              if (ct >= ct0) sRows[t * RSTR + 16 * ct + c] = tmp[ct];
        }
        __syncthreads();
        // S += Z * S[pivot rows, :], the panel's own column tile first: the
        // next panel reads it, and the other tiles' products run on the MFMA
        // pipe behind that panel's pivot searches
        double za[RT];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) za[rt] = sZ[g * RP + 16 * rt + c];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          if (ct < ct0) continue;
          const double b = sRows[g * RSTR + 16 * ct + c];
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) acc[rt][ct] = mfma64(za[rt], b, acc[rt][ct]);
        }
        __syncthreads();
      }
  }

  // ---- GK's A fragments for FE and every RHS tile (row 16 rta + l%16,
  // column of k-step s), in registers when they fit, loaded before FE needs them
  constexpr int KSR = S::GKREG ? S::KSP : 1;
  double gkf[KSR][TRH];
  int kend[TRH];
#pragma unroll
  for (int rta = 0; rta < TRH; ++rta) kend[rta] = (__builtin_amdgcn_readfirstlane(L.sTab[rta]) + 3) / 4;
  if constexpr (S::GKREG) {
    if (sparse) {
#pragma unroll
      for (int s = 0; s < KSR; ++s)
#pragma unroll
        for (int rta = 0; rta < TRH; ++rta) {
          const int k = L.sTab[TRH + rta * 4 * KSMAX + 4 * s + g];
          const int row = 16 * rta + c;
          if constexpr (TILE) gkf[s][rta] = s < kend[rta] ? GKe[(s * TRH + rta) * 64 + lane] : 0.0;
          else gkf[s][rta] = (s < kend[rta] && k != AE_NONE && row < NH) ? GKe[row + NH * k] : 0.0;
        }
    }
  }
  // ---- du: column DU of D', solution row k at pivot row p_k; its q rows,
  // DUDG and FE = Rh - GK x (schur_uq.c:146-186, 279-285); part 0 only
  if (part == 0) {
  if (c == DU % 16) {
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = sStep[16 * rt + 4 * i + g];
        if (k >= 0 && k < NU) sX[k] = acc[rt][CTD][i];
      }
  }
  __syncthreads();
  if constexpr (UQ) {
    for (int idx = lane; idx < NQ; idx += 64) {
      const int n = idx % NPV, m = (idx / NPV) % NCU, dd = idx / (NPV * NCU);
      double v = L.sMiRq[idx];
#pragma unroll
      for (int a = 0; a < NPV; ++a) v += L.sMiC[n + NPV * a + NPV * NPV * dd] * sX[a + NPV * m];
      sX[NU + idx] = v;
    }
    __syncthreads();
  }
  for (int r = lane; r < LDO; r += 64) {
    const double v = sX[r];
    if (!isfinite(v)) st |= 4;
    A.dudg[e * LDO + r] = v;
  }
  if constexpr (TILE) {
    // lane (g, c) sums row 16 rta + c over the k-steps' lane group g, then the
    // four lane groups are added
#pragma unroll
    for (int rta = 0; rta < TRH; ++rta) {
      // (the register fragments, or loads unrolled to the tile's bound so
      // they are in flight together)
      double v = 0.0;
#pragma unroll
      for (int s = 0; s < S::KSP; ++s) {
        if (s < kend[rta]) {
          const int k = L.sTab[TRH + rta * 4 * KSMAX + 4 * s + g];
          double gv;
          if constexpr (S::GKREG) gv = gkf[s < KSR ? s : 0][rta];
          else gv = GKe[(s * TRH + rta) * 64 + lane];
          if (k != AE_NONE) v += gv * sX[k];
        }
      }
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      const int row = 16 * rta + c;
      if (g == 0 && row < NH) {
        v = A.Rh[e * NH + row] - v;
        if (!isfinite(v)) st |= 4;
        A.fe[e * NH + row] = v;
      }
    }
  } else {
    for (int row = lane; row < NH; row += 64) {
      const int rta = row >> 4;
      const int klen = L.sTab[rta];
      const unsigned char* kt = L.sTab + TRH + rta * 4 * KSMAX;
      double v = 0.0;
#pragma unroll 8
      for (int idx = 0; idx < klen; ++idx) {
        const int k = kt[idx];
        v += GKe[row + NH * k] * sX[k];
      }
      v = A.Rh[e * NH + row] - v;
      if (!isfinite(v)) st |= 4;
      A.fe[e * NH + row] = v;
    }
  }
  }                                             // part 0

  // ---- A fragments of MiC for the q-row products (row n = l%16, k a = 4s + l/16)
  double mic[UQ ? ND : 1][S::KSV];
  if constexpr (UQ) {
#pragma unroll
    for (int dd = 0; dd < ND; ++dd)
#pragma unroll
      for (int s = 0; s < S::KSV; ++s) {
        const int a = 4 * s + g;
        mic[dd][s] = (c < NPV && a < NPV) ? L.sMiC[c + NPV * a + NPV * NPV * dd] : 0.0;
      }
  }
  __syncthreads();                            // sX (du) is reused below

  // ---- per RHS tile of this wavefront
#pragma unroll
  for (int tb = 0; tb < TRW; ++tb) {
    const int tbg = part * TRW + tb;           // the tile's place among all TRH
    if (tbg >= TRH) break;
    const int j = 16 * tbg + c;                // this lane's RHS column
    // u rows in pivot order
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = sStep[16 * rt + 4 * i + g];
        if (k >= 0 && k < NU) sX[k + NCOLP * c] = acc[rt][RT + tb][i];
      }
    // AE accumulators from H, block diagonal per face (on the sparse path;
    // the tile layout keeps only the face blocks)
    d4 ae[TRH];
#pragma unroll
    for (int rta = 0; rta < TRH; ++rta)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * rta + 4 * i + g;
        double v = 0.0;
        if (row < NH && j < NH) {
          const int t = s64_t<C>(j), fr = row / NB, fc = t / NB;
          if constexpr (TILE) {
            if (fr == fc) v = He[fr * NB * NB + (row - fr * NB) + NB * (t - fc * NB)];
          } else if (!sparse || fr == fc) {
            v = He[row + NH * t];
          }
        }
        ae[rta][i] = v;
      }
    __syncthreads();
    // q rows: dq(n, m, dd) = sum_a MiC_dd(n, a) du(a + NPV m) - MiE_dd(n, jf) [k == m]
    if constexpr (UQ) {
#pragma unroll
      for (int dd = 0; dd < ND; ++dd)
#pragma unroll
        for (int m = 0; m < NCU; ++m) {
          d4 q = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int s = 0; s < S::KSV; ++s) {
            const int a = 4 * s + g;
            const double b = a < NPV ? sX[a + NPV * m + NCOLP * c] : 0.0;
            q = mfma64(mic[dd][s], b, q);
          }
          const bool diag = j < NH && j / NDF == m;
          const int jf = j % NDF;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int n = 4 * i + g;
            if (n < NPV) {
              double v = q[i];
              if (diag) v -= L.sMiE[n + NPV * jf + NPV * NDF * dd];
              sX[NU + n + NPV * (m + NCU * dd) + NCOLP * c] = v;
            }
          }
        }
      __syncthreads();
    }
    // DUDG_DUH columns of this tile, coalesced over rows (two rows a lane
    // where the rows and the array allow 16-byte stores)
    if (LDO > 0) {
      double* Xe = A.dudg_duh + e * (int64_t)LDO * NH;
      if (LDO % 2 == 0 && (reinterpret_cast<uintptr_t>(A.dudg_duh) & 15) == 0) {
        for (int cc = 0; cc < 16; ++cc) {
          const int jj = 16 * tbg + cc;
          if (jj >= NH) break;
          for (int r2 = lane; 2 * r2 < LDO; r2 += 64) {
            double2 v;
            v.x = sX[2 * r2 + NCOLP * cc];
            v.y = sX[2 * r2 + 1 + NCOLP * cc];
            if (!isfinite(v.x) || !isfinite(v.y)) st |= 4;
            *reinterpret_cast<double2*>(Xe + 2 * r2 + (int64_t)LDO * s64_t<C>(jj)) = v;
          }
        }
      } else {
        for (int cc = 0; cc < 16; ++cc) {
          const int jj = 16 * tbg + cc;
          if (jj >= NH) break;
          for (int r = lane; r < LDO; r += 64) {
            const double v = sX[r + NCOLP * cc];
            if (!isfinite(v)) st |= 4;
            Xe[r + (int64_t)LDO * s64_t<C>(jj)] = v;
          }
        }
      }
    }
    // AE(:, tile) = H + GK X over the k-steps of each row tile's node set
    if (S::GKREG && sparse) {
#pragma unroll
      for (int s = 0; s < KSR; ++s)
#pragma unroll
        for (int rta = 0; rta < TRH; ++rta) {
          if (s >= kend[rta]) continue;
          const int k = L.sTab[TRH + rta * 4 * KSMAX + 4 * s + g];
          const double b = k != AE_NONE ? sX[k + NCOLP * c] : 0.0;
          ae[rta] = mfma64(gkf[s][rta], b, ae[rta]);
        }
    } else {
      int kmax = 0;
#pragma unroll
      for (int rta = 0; rta < TRH; ++rta) kmax = kend[rta] > kmax ? kend[rta] : kmax;
      constexpr int BATCH = TRH >= 4 ? 2 : 4;
#pragma unroll 1
      for (int s0 = 0; s0 < kmax; s0 += BATCH) {
        double a[BATCH][TRH], b[BATCH][TRH];
#pragma unroll
        for (int sb = 0; sb < BATCH; ++sb)
#pragma unroll
          for (int rta = 0; rta < TRH; ++rta) {
            const int s = s0 + sb;
            const int k = s < KSMAX ? L.sTab[TRH + rta * 4 * KSMAX + 4 * s + g] : AE_NONE;
            const int row = 16 * rta + c;
            const bool ok = s < kend[rta] && k != AE_NONE;
            if constexpr (TILE) a[sb][rta] = (s < kend[rta] && s < S::KSP) ? GKe[(s * TRH + rta) * 64 + lane] : 0.0;
            else a[sb][rta] = (ok && row < NH) ? GKe[row + NH * k] : 0.0;
            b[sb][rta] = ok ? sX[k + NCOLP * c] : 0.0;
          }
#pragma unroll
        for (int sb = 0; sb < BATCH; ++sb)
#pragma unroll
          for (int rta = 0; rta < TRH; ++rta)
            if (s0 + sb < kend[rta]) ae[rta] = mfma64(a[sb][rta], b[sb][rta], ae[rta]);
      }
    }
    __syncthreads();                          // X consumed; sX takes the AE tile
#pragma unroll
    for (int rta = 0; rta < TRH; ++rta)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * rta + 4 * i + g;
        if (row < NH) sX[row + NHP * c] = ae[rta][i];
      }
    __syncthreads();
    // whole AE columns: NH contiguous rows each (NH is a multiple of 4), two
    // rows a lane where the array allows 16-byte stores
    double* AEe = A.ae + e * (int64_t)NH * NH;
    if ((reinterpret_cast<uintptr_t>(A.ae) & 15) == 0) {
      for (int idx = lane; idx < 8 * NH; idx += 64) {
        const int cc = idx / (NH / 2), row = 2 * (idx - cc * (NH / 2));
        const int jj = 16 * tbg + cc;
        if (jj < NH) {
          double2 v;
          v.x = sX[row + NHP * cc];
          v.y = sX[row + 1 + NHP * cc];
          if (!isfinite(v.x) || !isfinite(v.y)) st |= 4;
          *reinterpret_cast<double2*>(AEe + row + NH * s64_t<C>(jj)) = v;
        }
      }
    } else {
      for (int idx = lane; idx < 16 * NH; idx += 64) {
        const int cc = idx / NH, row = idx - cc * NH;
        const int jj = 16 * tbg + cc;
        if (jj < NH) {
          const double v = sX[row + NHP * cc];
          if (!isfinite(v)) st |= 4;
          AEe[row + NH * s64_t<C>(jj)] = v;
        }
      }
    }
    __syncthreads();
  }
  st = wave_or(st);
  if (NW == 1) {
    if (lane == 0 && A.status) A.status[e] = st;
  } else if (lane == 0 && A.status && st) {
    atomicOr(&A.status[e], st);
  }
}

// blocks of schur64_kernel for ne elements: NW per element, the count rounded
// to whole rounds of the 8 XCDs
template <class C>
__host__ __device__ constexpr int64_t s64_blocks(int64_t ne) {
  constexpr int NW = S64Shape<C>::NW;
  return NW == 1 ? ne : 8 * ((ne + 7) / 8) * NW;
}

template <class C, bool TILE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(TILE ? S64Shape<C>::WPE : S64Shape<C>::WPE_REF,
                                                                   TILE ? S64Shape<C>::WPE : S64Shape<C>::WPE_REF))) void schur64_kernel(SchurArgs A) {
  __shared__ S64Lds<C> lds;
  constexpr int NW = S64Shape<C>::NW;
  if constexpr (NW == 1) {
    schur64_body<C, TILE>(A, blockIdx.x, 0, lds);
  } else {
    // workgroups go round-robin over the 8 XCDs: block b runs on XCD b % 8,
    // so an element's NW parts are blocks 8 k + x apart and share that XCD's
    // L2 for the D' they all read
    const int64_t b = blockIdx.x, x = b % 8, i = b / 8;
    schur64_body<C, TILE>(A, (i / NW) * 8 + x, (int)(i % NW), lds);
  }
}

// status words of ne elements cleared before a split elimination ORs into them
template <class C>
__global__ __launch_bounds__(256) void s64_zero_status(int32_t* st, int64_t ne) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < ne) st[i] = 0;
}

}  // namespace hdgx
