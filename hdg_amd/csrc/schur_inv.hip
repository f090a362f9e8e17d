// FP64 element elimination for u-only systems of four 16-row tiles (Euler
// tetrahedra p = 2, n_u = 50; 2D Euler p = 4, n_u = 60), kernel/schur_u.c:42-80
// with its mexFunction's AE and FE (:143-161), in two kernels:
//
//   inv64:   one wavefront per element.  Gauss-Jordan inversion in place with
//            partial pivoting, in panels of four columns held in MFMA
//            accumulator tiles [D | du], D = BD's u block.  When a panel's
//            pivot rows p_t are chosen, its columns are replaced by the
//            identity columns e_{p_t} they turn into, so after the last panel
//            column k holds column p_k of T = P D^-1 (P the pivot permutation)
//            and no second set of tiles is needed.  It writes D^-1 in natural
//            order over D in the element's workspace, du = D^-1 Ru (DUDG),
//            FE = Rh - GK du and the status word.
//   apply64: one wavefront per element and pair of RHS tiles, MFMA only:
//            X = D^-1 (-F) over k-steps of D^-1's columns, DUDG_DUH = X, and
//            AE = H + GK X.  X's accumulators are, register for register, the
//            B operand of the AE product (k-step 4 rt + i takes register i of
//            row tile rt), so X never leaves the registers.
//
// The pivot choice is the one of schur64.hip: the largest |.| of the column
// among rows not yet pivotal (to 2^-15 relative, piv_key), ties to the lowest
// row.  Against the other FP64
// forms the results agree to rounding.
#include <hip/hip_runtime.h>

#include "hdgx_device.h"
#include "hdgx_kernels.h"
#include "hdgx_shapes.h"

namespace hdgx {

template <class C>
struct InvShape {
  using S = S64Shape<C>;
  static constexpr bool USE = S::USE && S::RT == 4 && !C::UQ;
  static constexpr int NU = C::NU, NH = C::NH, RT = 4, R = 64, TRH = S::TRH;
  static constexpr int DU = NU, CTD = DU / 16;    // du's right-hand side in D's padding column
  static constexpr int KS = (NU + 3) / 4;         // k-steps over D^-1's columns (X's rows)
  static constexpr int TRB = 2;                   // RHS tiles per apply64 wavefront
  static constexpr int NWB = (TRH + TRB - 1) / TRB;
  static constexpr int RP = R + 1, RSTR = 16 * RT + 1;
  static constexpr int LDI = NU + 1;              // staging of T, odd leading dimension
  static constexpr int NSCR = 8 * RP + 4 * RSTR;
  // T is staged 32 of its columns at a time (two column tiles)
  static constexpr int NST = 32 * LDI > NSCR ? 32 * LDI : NSCR;
};

template <class C>
struct InvLds {
  using I = InvShape<C>;
  double sX[I::NST];              // panel, Z and pivot rows; then T for the output
  double sDu[I::R];
  int sStep[I::R];                // sStep[p] = k: pivot row p holds unknown k
  int sP[I::R];                   // sP[k] = p
};

// FB: GK in face blocks (FbShape, hdgx_shapes.h); perm the master's (npf, nfe)
// face-node table
template <class C, bool FB = false>
HDGX_DEV void inv64_body(const SchurArgs& A, const int64_t e, const int32_t* perm, InvLds<C>& L) {
  using I = InvShape<C>;
  constexpr int NU = I::NU, NH = I::NH, RT = I::RT, R = I::R, DU = I::DU, CTD = I::CTD;
  constexpr int RP = I::RP, RSTR = I::RSTR, LDI = I::LDI;
  const int lane = threadIdx.x;
  const int g = lane >> 4, c = lane & 15;
  if (e >= A.ne) return;
  double* const BDe = const_cast<double*>(A.BD) + e * (int64_t)NU * NU;
  int st = 0;

  d4 acc[RT][RT];
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
  double* const sPan = L.sX;
  double* const sZ = L.sX + 4 * RP;
  double* const sRows = L.sX + 8 * RP;
  L.sStep[lane] = -1;
  L.sP[lane] = -1;
  bool done = false;                           // row `lane` already pivotal
  __syncthreads();

#pragma unroll
  for (int ct0 = 0; ct0 < RT; ++ct0)
#pragma unroll 1
    for (int sub = 0; sub < 4; ++sub) {
      const int c0 = 4 * sub, k0 = 16 * ct0 + c0;
      if (k0 >= NU) break;
      const int np = NU - k0 < 4 ? NU - k0 : 4;
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
      for (int t = 0; t < 4; ++t) pr[t] = sPan[t * RP + lane];
      int pvt[4] = {-1, -1, -1, -1};
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (t >= np) break;
        unsigned key = 0;
        if (!done) key = piv_key(pr[t], lane);
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
        if (lane == 0) { L.sStep[p] = k0 + t; L.sP[k0 + t] = p; }
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) if (t < np && lane == pvt[t]) z[t] -= 1.0;
#pragma unroll
      for (int t = 0; t < 4; ++t) sZ[t * RP + lane] = z[t];
      // pre-panel pivot rows over every column (the earlier panels' columns
      // hold T's and take the update too); in the panel's own columns the
      // rows of the identity columns that replace them: delta(t, t')
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (t >= np) {
          if (g == 0)
#pragma unroll
            for (int ct = 0; ct < RT; ++ct) sRows[t * RSTR + 16 * ct + c] = 0.0;
          continue;
        }
        const int p = pvt[t];
        const int sel = 4 * (p >> 4) + ((p & 15) >> 2);
        double tmp[RT];
#pragma unroll
        for (int ct = 0; ct < RT; ++ct) tmp[ct] = 0.0;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (sel == 4 * rt + i)
#pragma unroll
              for (int ct = 0; ct < RT; ++ct) tmp[ct] = acc[rt][ct][i];
        if (g == (p & 3))
#pragma unroll
          for (int ct = 0; ct < RT; ++ct) {
            double v = tmp[ct];
            if (ct == ct0 && c >= c0 && c < c0 + np) v = (c - c0 == t) ? 1.0 : 0.0;
            sRows[t * RSTR + 16 * ct + c] = v;
          }
      }
      // the panel's columns become the identity columns e_{p_t'}
      {
        const int t2 = c - c0;
        if (t2 >= 0 && t2 < np) {
          const int pt = t2 == 0 ? pvt[0] : (t2 == 1 ? pvt[1] : (t2 == 2 ? pvt[2] : pvt[3]));
#pragma unroll
          for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[rt][ct0][i] = (16 * rt + 4 * i + g == pt) ? 1.0 : 0.0;
        }
      }
      __syncthreads();
      // S += Z * S[pivot rows, :], the panel's own column tile first (the
      // next panel reads it; the rest runs behind its pivot searches)
      double za[RT];
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) za[rt] = sZ[g * RP + 16 * rt + c];
#pragma unroll
      for (int k = 0; k < RT; ++k) {
        const int ct = (ct0 + k) % RT;
        const double b = sRows[g * RSTR + 16 * ct + c];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) acc[rt][ct] = mfma64(za[rt], b, acc[rt][ct]);
      }
      __syncthreads();
    }

  // ---- du (column DU: unknown k at pivot row p_k) and T to LDS
  if (c == DU % 16) {
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = L.sStep[16 * rt + 4 * i + g];
        if (k >= 0 && k < NU) L.sDu[k] = acc[rt][CTD][i];
      }
  }
  // D^-1(a, b) = T(p_a, column holding T's column b) = T(p_a, sStep[b]),
  // written over D; T staged in LDS 32 columns at a time
  // (a singular D, status bit 2, can leave a row unpivoted or pivot a padding
  // row: those entries become NaN rather than read past T)
  double* const sT = L.sX;                     // T(r, 32 h + k) at r + LDI k
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int ct = 2 * h; ct < 2 * h + 2; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = 16 * rt + 4 * i + g, col = 16 * ct + c;
          if (row < NU && col < NU) sT[row + LDI * (col - 32 * h)] = acc[rt][ct][i];
        }
    __syncthreads();
    for (int idx = lane; idx < NU * NU; idx += 64) {
      const int a = idx % NU, b = idx / NU;
      const int pa = L.sP[a], kb = L.sStep[b];
      if ((h == 0) != (kb < 32)) continue;     // the other half's column (kb < 0: the first)
      BDe[idx] = (pa >= 0 && pa < NU && kb >= 0) ? sT[pa + LDI * (kb - 32 * h)] : __builtin_nan("");
    }
    __syncthreads();
  }
  const int LDO = C::NPV * A.nco;
  for (int r = lane; r < LDO; r += 64) {
    const double v = L.sDu[r];
    if (!isfinite(v)) st |= 4;
    A.dudg[e * LDO + r] = v;
  }
  if constexpr (FB) {
    using F = FbShape<C>;
    constexpr int NB = F::NB, NPF = C::NPF, NPV = C::NPV;
    const double* GKe = A.GK + e * (int64_t)F::BLK;
    for (int row = lane; row < NH; row += 64) {
      const int f = row / NB, r = row - f * NB;
      const double* G = GKe + (int64_t)NB * NB * f;
      double v = 0.0;
#pragma unroll 6
      for (int q = 0; q < NB; ++q) v += G[r + NB * q] * L.sDu[perm[q % NPF + NPF * f] + NPV * (q / NPF)];
      v = A.Rh[e * NH + row] - v;
      if (!isfinite(v)) st |= 4;
      A.fe[e * NH + row] = v;
    }
  } else {
    const double* GKe = A.GK + e * (int64_t)NH * NU;
    for (int row = lane; row < NH; row += 64) {
      double v = 0.0;
#pragma unroll 10
      for (int k = 0; k < NU; ++k) v += GKe[row + NH * k] * L.sDu[k];
      v = A.Rh[e * NH + row] - v;
      if (!isfinite(v)) st |= 4;
      A.fe[e * NH + row] = v;
    }
  }
  st = wave_or(st);
  if (lane == 0 && A.status) A.status[e] = st;
}

// part w of element e: RHS tiles TRB w .. TRB w + TRB - 1
template <class C>
HDGX_DEV void apply64_body(const SchurArgs& A, const int64_t e, const int part) {
  using I = InvShape<C>;
  constexpr int NU = I::NU, NH = I::NH, RT = I::RT, KS = I::KS, TRB = I::TRB, TRH = I::TRH;
  const int lane = threadIdx.x;
  const int g = lane >> 4, c = lane & 15;
  if (e >= A.ne) return;
  const double* Di = A.BD + e * (int64_t)NU * NU;
  const double* Fe = A.F + e * (int64_t)NU * NH;
  const double* GKe = A.GK + e * (int64_t)NH * NU;
  const double* He = A.H + e * (int64_t)NH * NH;
  int st = 0;
  int jcol[TRB];                               // this lane's trace column per tile (-1 past NH)
#pragma unroll
  for (int tb = 0; tb < TRB; ++tb) {
    const int j = 16 * (part * TRB + tb) + c;
    jcol[tb] = j < NH ? s64_t<C>(j) : -1;
  }

  // ---- X = D^-1 (-F)
  d4 x[RT][TRB];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int tb = 0; tb < TRB; ++tb) x[rt][tb] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int k = 4 * s + g;
    double a[RT], b[TRB];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) a[rt] = (16 * rt + c < NU && k < NU) ? Di[16 * rt + c + NU * k] : 0.0;
#pragma unroll
    for (int tb = 0; tb < TRB; ++tb) b[tb] = (k < NU && jcol[tb] >= 0) ? -Fe[k + NU * jcol[tb]] : 0.0;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int tb = 0; tb < TRB; ++tb) x[rt][tb] = mfma64(a[rt], b[tb], x[rt][tb]);
  }

  // ---- DUDG_DUH columns of these tiles (u rows kept)
  const int LDO = C::NPV * A.nco;
  if (LDO > 0) {
    double* Xe = A.dudg_duh + e * (int64_t)LDO * NH;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int tb = 0; tb < TRB; ++tb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = 16 * rt + 4 * i + g;
          if (row < LDO && jcol[tb] >= 0) {
            const double v = x[rt][tb][i];
            if (!isfinite(v)) st |= 4;
            Xe[row + (int64_t)LDO * jcol[tb]] = v;
          }
        }
  }

  // ---- AE(:, tiles) = H + GK X, one row tile of AE at a time; the next row
  // tile's GK and H are loaded before this one's products and stores (which
  // would otherwise hold the loads back behind them)
  double* AEe = A.ae + e * (int64_t)NH * NH;
  auto load_rt = [&](int rta, double (&gk)[KS], double (&hv)[TRB][4]) {
    const int rowa = 16 * rta + c;              // GK row of this lane's A operand
#pragma unroll
    for (int s = 0; s < KS; ++s) gk[s] = (rowa < NH && 4 * s + g < NU) ? GKe[rowa + NH * (4 * s + g)] : 0.0;
#pragma unroll
    for (int tb = 0; tb < TRB; ++tb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * rta + 4 * i + g;
        hv[tb][i] = (row < NH && jcol[tb] >= 0) ? He[row + NH * jcol[tb]] : 0.0;
      }
  };
  double gk[KS], gkn[KS], hv[TRB][4], hvn[TRB][4];
  load_rt(0, gk, hv);
#pragma unroll
  for (int rta = 0; rta < TRH; ++rta) {
    if (rta + 1 < TRH) load_rt(rta + 1, gkn, hvn);
    d4 ae[TRB];
#pragma unroll
    for (int tb = 0; tb < TRB; ++tb)
#pragma unroll
      for (int i = 0; i < 4; ++i) ae[tb][i] = hv[tb][i];
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int tb = 0; tb < TRB; ++tb) ae[tb] = mfma64(gk[s], x[s >> 2][tb][s & 3], ae[tb]);
#pragma unroll
    for (int tb = 0; tb < TRB; ++tb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * rta + 4 * i + g;
        if (row < NH && jcol[tb] >= 0) {
          const double v = ae[tb][i];
          if (!isfinite(v)) st |= 4;
          AEe[row + (int64_t)NH * jcol[tb]] = v;
        }
      }
#pragma unroll
    for (int s = 0; s < KS; ++s) gk[s] = gkn[s];
#pragma unroll
    for (int tb = 0; tb < TRB; ++tb)
#pragma unroll
      for (int i = 0; i < 4; ++i) hv[tb][i] = hvn[tb][i];
  }
  st = wave_or(st);
  if (lane == 0 && A.status && st) atomicOr(&A.status[e], st);
}

// face-block form of apply64 (FbShape): wavefront `part` = face fc takes the
// RHS tiles of that face's NB trace columns.  X(:, face fc) = D^-1(:, u rows
// of fc) (-F_fc) over face-local k-steps; X goes to LDS once, and AE's row
// tiles of face fr take GK_fr against the rows of X on face fr.
template <class C>
struct ApplyFbLds {
  static constexpr int XS = 64 + 1;               // X column stride
  double sX[FbShape<C>::NBT * 16 * XS];
  int sU[C::NFE * FbShape<C>::NB];                // u row of face-local q on face f at f NB + q
};

template <class C>
HDGX_DEV void apply64_fb_body(const SchurArgs& A, const int64_t e, const int fc, const int32_t* perm,
                              ApplyFbLds<C>& L) {
  using I = InvShape<C>;
  using F = FbShape<C>;
  constexpr int NU = I::NU, NH = I::NH, RT = I::RT, NB = F::NB, NBT = F::NBT, KSF = F::KSF;
  constexpr int NFE = C::NFE, NPF = C::NPF, NPV = C::NPV, XS = ApplyFbLds<C>::XS;
  const int lane = threadIdx.x;
  const int g = lane >> 4, c = lane & 15;
  if (e >= A.ne) return;
  for (int i = lane; i < NFE * NB; i += 64) {
    const int f = i / NB, q = i - f * NB;
    L.sU[i] = perm[q % NPF + NPF * f] + NPV * (q / NPF);
  }
  __syncthreads();
  const double* Di = A.BD + e * (int64_t)NU * NU;
  const double* Fb = A.F + e * (int64_t)F::BLK + (int64_t)NB * NB * fc;
  const double* GKe = A.GK + e * (int64_t)F::BLK;
  const double* Hb = A.H + e * (int64_t)F::BLK + (int64_t)NB * NB * fc;
  int st = 0;
  // GK_fr's A operands for face fr (this lane: row 16 sr + c, column 4 s + g),
  // loaded one face ahead of their products (the AE stores would otherwise
  // hold every load back behind them); face 0's and H_fc's before X
  auto load_gk = [&](int fr, double (&v)[NBT][KSF]) {
    const double* G = GKe + (int64_t)NB * NB * fr;
#pragma unroll
    for (int sr = 0; sr < NBT; ++sr)
#pragma unroll
      for (int s = 0; s < KSF; ++s) {
        const int ra = 16 * sr + c, q = 4 * s + g;
        v[sr][s] = (ra < NB && q < NB) ? G[ra + NB * q] : 0.0;
      }
  };
  double gkc[NBT][KSF], gkn[NBT][KSF];
  load_gk(0, gkc);
  // X and the AE blocks are formed transposed (operands swapped), so register
  // i of lane (g, c) holds row 16 t + c, column 16 tb + 4 i + g of its tile:
  // sixteen lanes store sixteen consecutive rows of a column
  double hv[NBT][NBT][4];
#pragma unroll
  for (int sr = 0; sr < NBT; ++sr)
#pragma unroll
    for (int tb = 0; tb < NBT; ++tb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int rr = 16 * sr + c, r = 16 * tb + 4 * i + g;
        hv[sr][tb][i] = (rr < NB && r < NB) ? Hb[rr + NB * r] : 0.0;
      }

  // ---- X = D^-1(:, u rows of fc) (-F_fc), face-local k (X^T = (-F_fc)^T D^-T)
  d4 x[RT][NBT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int tb = 0; tb < NBT; ++tb) x[rt][tb] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < KSF; ++s) {
    const int q = 4 * s + g;
    const int uq = q < NB ? L.sU[fc * NB + q] : 0;
    double a[RT], b[NBT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) a[rt] = (16 * rt + c < NU && q < NB) ? Di[16 * rt + c + NU * uq] : 0.0;
#pragma unroll
    for (int tb = 0; tb < NBT; ++tb) {
      const int r = 16 * tb + c;
      b[tb] = (q < NB && r < NB) ? -Fb[q + NB * r] : 0.0;
    }
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int tb = 0; tb < NBT; ++tb) x[rt][tb] = mfma64(b[tb], a[rt], x[rt][tb]);
  }

  // ---- DUDG_DUH columns of face fc, and X to LDS
  const int LDO = C::NPV * A.nco;
  double* Xe = A.dudg_duh + e * (int64_t)LDO * NH;
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int tb = 0; tb < NBT; ++tb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * rt + c, r = 16 * tb + 4 * i + g;
        const double v = x[rt][tb][i];
        L.sX[row + XS * r] = v;
        if (row < LDO && r < NB) {
          if (!isfinite(v)) st |= 4;
          Xe[row + (int64_t)LDO * (fc * NB + r)] = v;
        }
      }
  __syncthreads();

  // ---- AE(rows of face fr, columns of face fc) = [fr == fc] H_fc + GK_fr X(u rows of fr, :)
  double* AEe = A.ae + e * (int64_t)NH * NH;
#pragma unroll
  for (int fr = 0; fr < NFE; ++fr) {
    if (fr + 1 < NFE) load_gk(fr + 1, gkn);
#pragma unroll
    for (int sr = 0; sr < NBT; ++sr) {
      d4 ae[NBT];
#pragma unroll
      for (int tb = 0; tb < NBT; ++tb)
#pragma unroll
        for (int i = 0; i < 4; ++i) ae[tb][i] = fr == fc ? hv[sr][tb][i] : 0.0;
#pragma unroll
      for (int s = 0; s < KSF; ++s) {
        const int q = 4 * s + g;
        const double a = gkc[sr][s];
        const int uq = q < NB ? L.sU[fr * NB + q] : 0;
#pragma unroll
        for (int tb = 0; tb < NBT; ++tb) {
          const double b = q < NB ? L.sX[uq + XS * (16 * tb + c)] : 0.0;
          ae[tb] = mfma64(b, a, ae[tb]);
        }
      }
#pragma unroll
      for (int tb = 0; tb < NBT; ++tb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int rr = 16 * sr + c, r = 16 * tb + 4 * i + g;
          if (rr < NB && r < NB) {
            const double v = ae[tb][i];
            if (!isfinite(v)) st |= 4;
            AEe[(fr * NB + rr) + (int64_t)NH * (fc * NB + r)] = v;
          }
        }
    }
#pragma unroll
    for (int sr = 0; sr < NBT; ++sr)
#pragma unroll
      for (int s = 0; s < KSF; ++s) gkc[sr][s] = gkn[sr][s];
  }
  st = wave_or(st);
  if (lane == 0 && A.status && st) atomicOr(&A.status[e], st);
}

template <class C, bool FB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3))) void inv64_kernel(SchurArgs A,
                                                                                             const int32_t* perm) {
  __shared__ InvLds<C> lds;
  inv64_body<C, FB>(A, blockIdx.x, perm, lds);
}

// blocks of apply64_kernel for ne elements: NWB per element, whole rounds of
// the 8 XCDs (block b runs on XCD b % 8; an element's parts share it)
template <class C, bool FB = false>
__host__ __device__ constexpr int64_t apply64_blocks(int64_t ne) {
  return 8 * ((ne + 7) / 8) * (FB ? C::NFE : InvShape<C>::NWB);
}

template <class C>
__global__ __launch_bounds__(64) void apply64_kernel(SchurArgs A) {
  constexpr int NW = InvShape<C>::NWB;
  const int64_t b = blockIdx.x, x = b % 8, i = b / 8;
  apply64_body<C>(A, (i / NW) * 8 + x, (int)(i % NW));
}

template <class C>
__global__ __launch_bounds__(64) void apply64_fb_kernel(SchurArgs A, const int32_t* perm) {
  if constexpr (FbShape<C>::USE) {
    constexpr int NW = C::NFE;
    __shared__ ApplyFbLds<C> lds;
    const int64_t b = blockIdx.x, x = b % 8, i = b / 8;
    apply64_fb_body<C>(A, (i / NW) * 8 + x, (int)(i % NW), perm, lds);
  }
}

}  // namespace hdgx
