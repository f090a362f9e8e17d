// Kernel configurations and argument blocks shared by the hdgx kernels and
// the host dispatch.
#pragma once

#include <stdint.h>

namespace hdgx {

constexpr int tri_npts(int pg) { return pg <= 1 ? 1 : pg == 2 ? 3 : pg == 3 ? 4 : pg == 4 ? 6 : pg == 5 ? 7 : pg == 6 ? 12 : pg == 7 ? 13 : 16; }
constexpr int tet_npts(int pg) { return pg <= 2 ? 4 : pg == 3 ? 5 : pg == 4 ? 11 : pg == 5 ? 15 : 24; }

// Element configuration fixed at compile time: physics, dimension and order.
// Quadrature is the reference default pgauss = 2p (master/mkmaster.m:10-12).
template <int PHYS_, int ND_, int P_>
struct Cfg {
  static constexpr int PHYS = PHYS_, ND = ND_, P = P_;
  static constexpr int NPV = (ND == 2) ? (P + 1) * (P + 2) / 2 : (P + 1) * (P + 2) * (P + 3) / 6;
  static constexpr int NPF = (ND == 2) ? P + 1 : (P + 1) * (P + 2) / 2;
  static constexpr int NFE = ND + 1;
  static constexpr int NDF = NPF * NFE;
  static constexpr int NGV = (ND == 2) ? tri_npts(2 * P) : tet_npts(2 * P);
  static constexpr int NGF = (ND == 2) ? P + 1 : tri_npts(2 * P);
  static constexpr int NCU = ND + 2;
  static constexpr int NC = (PHYS == 1) ? NCU * (ND + 1) : NCU;
  static constexpr bool UQ = (PHYS == 1);
  static constexpr int NU = NPV * NCU;       // local u unknowns
  static constexpr int NH = NCU * NDF;       // trace unknowns
  static constexpr int NCOL = NPV * NC;      // columns of BD / GK, rows of DUDG_DUH
  static constexpr int NR = NGV * (ND + 1);  // volume quadrature rows (value + derivatives)
};

// Master element tables on the device (column-major, reference shapes).
struct DevMaster {
  const double* shapvt;           // (ngv, npv, nd+1)
  const double* shapvg;           // (npv, ngv, nd+1)
  const double* svv;              // shapvgdotshapvl (npv*npv, ngv, nd+1)
  const double* shapft;           // (ngf, npf, nd)
  const double* shapfg;           // (npf, ngf, nd)
  const double* sff;              // shapfgdotshapfc (npf*npf, ngf, nd)
  const double* shapmv;           // (ngv, npv, nd+1)
  const double* shapmf;           // (ngf, npf, nd)
  // B fragments of the quadrature's BD product (BdShape): per row tile R of
  // volume node pairs (jj, ii) (row jj + npv ii) and k-step s, 64 doubles,
  // lane l = SV(row 4 s + l / 16, pair 16 R + l % 16); SV rows r < ngv (nd+1)
  // are shapvgdotshapvl, rows ngv (nd+1) + j ngf + g shapfgdotshapfc(loc(jj, j),
  // loc(ii, j), g) where both nodes lie on face j, else 0
  const double* bdv;
  const int32_t* perm;            // (npf, nfe)
  const int32_t* permgeom;        // (npf, nfe)
  const int32_t* loc;             // (npv, nfe): position of volume node v on face j, or -1
  // AE k-step tables of the face-sparse GK (u only: [0], u and q: [1]): per
  // AE row tile of 16 rows its count of nonzero GK columns, then per tile
  // 4*KSMAX column indices (255 past the count), KSMAX = ceil(max(npv,
  // 2 npf) * nc / 4)
  const unsigned char* aek[2];
};

// Physics parameters passed by value.
struct DevApp {
  double gam, eps, Re, Pr, tau;
  double Minf;                     // NS free-stream Mach number (isothermal wall)
  double fc_u, fc_q;
  int32_t tdep, av_col, ncd, nbc;
  int32_t bcm[8];
  double bcs[8 * 5];               // bcs[i*5 + c]
};

struct QuadArgs {
  int64_t ne;
  const double* dgnodes;
  const int32_t* bf;
  const double* udg;
  const double* uh;
  const double* sh;
  double *BD, *F, *GK, *H, *Ru, *Rh, *M, *C, *E, *Rq;
  int32_t* status;
};

// q recovery (getq): per element q = (fc_q M)^-1 (C u - E uh) + SH_q/fc_q.
struct GetqArgs {
  int64_t ne;
  int32_t nc, ncd;                 // columns of udg / sh and of dgnodes
  double fc_q;
  const double* dgnodes;           // (npv, ncd, ne)
  const double* udg;               // (npv, nc, ne): the first ncu columns are read
  const double* uh;                // (ncu, npf*nfe, ne)
  const double* sh;                // (npv, nc, ne) or null: its last ncu*nd columns are added / fc_q
  double* qdg;                     // element e at qdg + e*ldq: (npv, ncu*nd)
  int64_t ldq;
  const double* shapnv;            // (npv, npv, nd+1) nodal basis at the nodes, or null
  double* qq;                      // (npv, ncu*nd, ne) or null
};

// HDGX_INPLACE: the MEX files' side effects on their inputs (schur_uq.c,
// schur_u.c); uq = 0 for schur_u (only BD is written).
struct InplaceArgs {
  int64_t ne;
  int32_t npv, ncu, nd, ndf, uq;
  double *M, *C, *E, *BD, *F, *Rq, *Ru;
};

struct SchurArgs {
  int64_t ne;
  const double *M, *C, *E, *BD, *F, *GK, *H, *Rq, *Ru, *Rh;
  double *dudg, *dudg_duh, *ae, *fe;
  int32_t* status;
  int32_t prec;                    // 0 fp64, 1 fp32 mfma + fp64 refinement
  int32_t nco;                     // DUDG / DUDG_DUH rows kept per node: nc, ncu or 0
  const unsigned char* aek;        // AE k-step table (DevMaster::aek) when GK is face sparse, else null
  int32_t bd_scratch;              // 1: BD is the library's workspace and its D block may be overwritten
};

// One request's element range inside a batched launch: its arrays and the
// index of its physics parameters in the launch's DevApp table.
struct ElemSeg {
  int64_t e0, ne;                  // first element in the batch numbering, count
  const double* dgnodes;
  const int32_t* bf;
  const double* udg;
  const double* uh;
  const double* sh;
  double *ae, *fe, *dudg, *dudg_duh;
  int32_t* status;
  int32_t app;
  int32_t order;                   // polynomial order p (mixed-order launches)
};

// Segment table of one launch over batch elements g0 .. g0+total-1; the
// element matrices between the two kernels live in the workspace arrays,
// indexed by g - g0.
struct SegTable {
  const ElemSeg* segs;
  const DevApp* apps;
  int32_t nseg;
  int64_t g0, total;
  double *BD, *F, *GK, *H, *Ru, *Rh, *M, *C, *E, *Rq;
};

// One persistent launch over the elements of several orders p (one physics
// and dimension): batch elements are ordered by decreasing p; order p owns
// batch elements first[p] .. end[p]-1 and its own workspace arrays.
constexpr int HDGX_MAXP = 5;
struct MixedTable {
  const ElemSeg* segs;
  const DevApp* apps;
  int32_t nseg;
  int64_t total;
  int64_t first[HDGX_MAXP + 1], end[HDGX_MAXP + 1];
  DevMaster masters[HDGX_MAXP];
  double *BD[HDGX_MAXP], *F[HDGX_MAXP], *GK[HDGX_MAXP], *H[HDGX_MAXP], *Ru[HDGX_MAXP], *Rh[HDGX_MAXP];
  double *M[HDGX_MAXP], *C[HDGX_MAXP], *E[HDGX_MAXP], *Rq[HDGX_MAXP];
};

}  // namespace hdgx
