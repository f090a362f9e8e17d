// Compile-time shapes shared by the kernels and the host dispatch: the FP64
// elimination for one to three row tiles (schur64.hip) and the layout of the
// element-matrix workspace between the quadrature and elimination kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "hdgx_kernels.h"

namespace hdgx {

// The quadrature's BD product: BD^T rows (c, k) (NCC = ncu nc, CTB tiles),
// columns the volume node pairs (jj, ii) (NVV = npv^2, RTB tiles), k over the
// NR volume rows and the nfe ngf face points (KT, KSB k-steps).  The host's
// table of B fragments (bd_table) uses the same sizes.
template <class C>
struct BdShape {
  static constexpr int NVV = C::NPV * C::NPV, RTB = (NVV + 15) / 16;
  static constexpr int KT = C::NR + C::NFE * C::NGF, KSB = (KT + 3) / 4;
  static constexpr int NCC = C::NCU * C::NC, CTB = (NCC + 15) / 16;
};

template <class C>
struct S64Shape {
  static constexpr int NU = C::NU, NH = C::NH, NPV = C::NPV, NCU = C::NCU, ND = C::ND, NC = C::NC;
  static constexpr int NDF = C::NDF, NPF = C::NPF, NFE = C::NFE, NCOL = C::NCOL;
  static constexpr int RT = (NU + 15) / 16, R = 16 * RT;
  static constexpr int DU = NU;                  // padding column with du's right-hand side
  static constexpr int CTD = DU / 16;            // its tile
  static constexpr int TRH = (NH + 15) / 16;     // RHS tiles = AE row tiles
  static constexpr int CT = RT + TRH;            // column tiles of the system
  // four row tiles: the RHS tiles are split over NW one-wavefront blocks
  // (each repeats the elimination of D' and carries TRW of them)
  static constexpr int NW = RT == 4 ? (TRH >= 8 ? 4 : 2) : 1;
  static constexpr int TRW = (TRH + NW - 1) / NW;
  static constexpr int CTW = RT + TRW;          // column tiles one wavefront holds
  // one to three row tiles with n_u a multiple of 4, or four row tiles of a
  // u-only system (Euler p = 2 tetrahedra, n_u = 50; p = 4 triangles, 60) at
  // one wavefront per SIMD, the accumulators spread over VGPRs and AGPRs
  static constexpr bool USE = NH % 4 == 0 && NU < R &&
                              ((RT <= 3 && NU % 4 == 0) || (RT == 4 && !C::UQ));
  static constexpr int KB = NPV * NCU * ND;      // columns of B (u,q)
  static constexpr int KBS = (KB + 3) / 4;
  static constexpr int KSV = (NPV + 3) / 4;      // k-steps of the q-row products
  // AE k-step table (same constants as the host's ae_table)
  static constexpr int NNMAX = NPV > 2 * NPF ? NPV : 2 * NPF;
  static constexpr int KSMAX = (NNMAX * NC + 3) / 4;
  // k-steps of a face-sparse AE row tile at most: its rows meet faces
  // j0..j1, so at most min(NPV, faces * NPF) nodes
  static constexpr int ksp_tile(int rta) {
    const int j0 = (16 * rta) / (NCU * NPF);
    int j1 = (16 * rta + 15) / (NCU * NPF);
    if (j1 > NFE - 1) j1 = NFE - 1;
    int nn = (j1 - j0 + 1) * NPF;
    if (nn > NPV) nn = NPV;
    return (nn * NC + 3) / 4;
  }
  static constexpr int ksp() {
    int m = 0;
    for (int r = 0; r < TRH; ++r) m = ksp_tile(r) > m ? ksp_tile(r) : m;
    return m;
  }
  static constexpr int KSP = ksp();
  // GK's A fragments stay in registers across the RHS tiles when they fit
  // (with three row tiles in the registers D' frees after the elimination)
  static constexpr bool GKREG = KSP * TRH <= (RT == 3 ? 36 : 16);
  // wavefronts per SIMD the elimination is compiled for, as many as the
  // registers allow without spills: the tile-layout kernel (WPE), the batched
  // one (WPE_SEG) and the reference-layout one (WPE_REF); u-only systems of
  // one row tile (2D Euler p = 1) 8, 6 and 4; four for u-only two-row-tile
  // systems with at most three RHS tiles (2D Euler p = 2)
  static constexpr int WPE = RT == 1 ? (C::UQ ? 4 : 8) : (RT == 2 ? (!C::UQ && TRH <= 3 ? 4 : 3) : 2);
  static constexpr int WPE_SEG = RT == 1 && !C::UQ ? 6 : WPE;
  static constexpr int WPE_REF = RT == 1 ? 4 : WPE;
};

// Face-block form of the u-only four-row-tile systems (Euler tetrahedra p = 2:
// 30 trace rows per face, two tiles; 2D Euler p = 4: 20 rows, two tiles):
// F, GK and H are nonzero only where a trace row or column and a volume node
// share a face (faceint.m:316-359; H block diagonal per face), so each is kept
// as NFE blocks of NB x NB, NB = NCU * NPF.  Face-local indices: trace r =
// k + NCU b (component k, face node b: trace column f NB + r), volume
// q = a + NPF m (face node a, i.e. volume node perm(a, f), component m: u row
// perm(a, f) + NPV m).
template <class C>
struct FbShape {
  using S = S64Shape<C>;
  static constexpr int NB = C::NCU * C::NPF;
  static constexpr int NBT = (NB + 15) / 16;     // 16-row tiles per face
  static constexpr int KSF = (NB + 3) / 4;       // face-local k-steps
  // where the blocks hold fewer words than the dense arrays (Euler tetrahedra
  // p = 2: 3 600 of 6 000; 2D Euler p = 4: 1 200 of 3 600)
  static constexpr bool USE = S::USE && S::RT == 4 && !C::UQ && C::NFE * NB * NB < C::NH * C::NU;
  static constexpr size_t BLK = (size_t)C::NFE * NB * NB;   // words of F, GK or H
};

// Per-element words of the element-matrix workspace in hdgx_elemental and the
// batched launches.  The reference layout keeps every array as
// kernel/schur_primal.m receives it; the tile layout (FP64 configurations of
// schur64) stores F and H as face blocks and GK in the elimination's MFMA
// fragment order, one 64-lane vector per (tile, register) so that every GK
// load of the elimination reads 512 contiguous bytes (NB = NCU NPF):
//   F  NFE blocks of NB x NB (F couples a face's trace nodes only with the
//      volume nodes on that face, faceint.m:316-327): block f holds F(u row
//      of q on face f, trace column f NB + r) at q + NB r, with q = a + NPF m
//      (face node a, i.e. volume node perm(a, f), component m) and r = k +
//      NCU b (component k, face node b);
//   GK (KSP, TRH) vectors: lane 16g + c of vector (s, rta) = GK(row 16 rta + c,
//      column of k-step s, lane group g of the AE table), only the face-sparse
//      columns (faceint.m:354-358);
//   H  NFE blocks of NB x NB (H is block diagonal per face,
//      faceint.m:347-359): block f holds H(f NB + r, f NB + r') at r + NB r'.
// For FbShape configurations GK is face blocks too: block f: (r, q) at r + NB q.
// The other arrays (BD, Ru, Rh, M, C, E, Rq) keep the reference layout.
template <class C>
struct WsLayout {
  using S = S64Shape<C>;
  static constexpr bool FB = FbShape<C>::USE;
  static constexpr bool TILE_OK = S::USE && (S::RT <= 3 || FB);   // fragment tiles or face blocks
  static constexpr size_t NU = C::NU, NH = C::NH, NCOL = C::NCOL, NPV = C::NPV, ND = C::ND;
  static constexpr size_t bd = NU * NCOL, ru = NU, rh = NH;
  static constexpr size_t m = C::UQ ? NPV * NPV : 0, cm = C::UQ ? NPV * NPV * ND : 0;
  static constexpr size_t em = C::UQ ? NPV * C::NDF * ND : 0, rq = C::UQ ? NPV * C::NCU * ND : 0;
  static constexpr size_t fblk = (size_t)C::NFE * (C::NCU * C::NPF) * (C::NCU * C::NPF);
  static constexpr size_t f(bool tile) { return tile ? fblk : NU * NH; }
  static constexpr size_t gk(bool tile) {
    return tile ? (FB ? FbShape<C>::BLK : (size_t)S::KSP * S::TRH * 64) : NH * NCOL;
  }
  static constexpr size_t hblk = (size_t)C::NFE * (C::NCU * C::NPF) * (C::NCU * C::NPF);
  static constexpr size_t h(bool tile) { return tile ? hblk : NH * NH; }
  static constexpr size_t words(bool tile) { return bd + f(tile) + gk(tile) + h(tile) + ru + rh + m + cm + em + rq; }
  static constexpr size_t max_words() {
    return TILE_OK && words(true) > words(false) ? words(true) : words(false);
  }
};

// the workspace of n elements carved into its arrays (each element-major)
struct WsPtrs {
  double *BD, *F, *GK, *H, *Ru, *Rh, *M, *C, *E, *Rq;
};
template <class C>
__host__ __device__ inline WsPtrs ws_carve(double* w, int64_t n, bool tile) {
  using L = WsLayout<C>;
  WsPtrs p{};
  p.BD = w; w += L::bd * n;
  p.F = w; w += L::f(tile) * n;
  p.GK = w; w += L::gk(tile) * n;
  p.H = w; w += L::h(tile) * n;
  p.Ru = w; w += L::ru * n;
  p.Rh = w; w += L::rh * n;
  if (C::UQ) {
    p.M = w; w += L::m * n;
    p.C = w; w += L::cm * n;
    p.E = w; w += L::em * n;
    p.Rq = w;
  }
  return p;
}

}  // namespace hdgx
