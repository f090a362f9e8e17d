/*
 * hdgx -- MI355X (gfx950) HDG elemental hot path behind a C ABI.
 *
 * Every array is column-major with the reference's MATLAB index order, so a
 * reference caller hands over exactly the arrays it already holds.  Array
 * arguments of the compute entry points are device (HBM) pointers; the master
 * and app structs and the tables they point to live in host memory and are
 * copied to the device by the library (cached per context).
 *
 * Sizes: nd spatial dim, npv nodes/elem, npf nodes/face, nfe faces/elem,
 * ndf = npf*nfe, ncu state comps, nc = ncu (u only) or ncu*(nd+1) (u and q),
 * ncd dgnodes columns (nd, or nd+1 with an artificial-viscosity column).
 */
#ifndef HDGX_H
#define HDGX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
enum {
  HDGX_OK = 0,
  HDGX_ERR_ARG = 1,         /* null pointer, bad size, inconsistent dims */
  HDGX_ERR_UNSUPPORTED = 2, /* no kernel instantiated for this (physics, nd, p) */
  HDGX_ERR_DEVICE = 3,      /* HIP runtime error */
  HDGX_ERR_NOMEM = 4
};

enum { HDGX_PHYS_EULER = 0, HDGX_PHYS_NS = 1 };

/* arithmetic of the local elimination (see hdgx_set_precision) */
enum { HDGX_PREC_FP64 = 0, HDGX_PREC_FP32_IR = 1 };

/* per-element status bits written to status[e] (0 = clean).  The reference
 * ignores LAPACK's info (kernel/schur_uq.c:69,136); these report it instead. */
enum {
  HDGX_ST_NONSPD_M = 1,   /* mass matrix not SPD (Cholesky pivot <= 0) */
  HDGX_ST_SINGULAR_D = 2, /* zero pivot in the elimination of D */
  HDGX_ST_NONFINITE = 4,  /* NaN/Inf in an output of the element */
  HDGX_ST_IR_SLOW = 8     /* FP32+refinement did not reach its residual target */
};

typedef struct hdgx_ctx hdgx_ctx;

/* Master element tables: mkmaster + preprocess products
 * (master/mkmaster.m:1-96, util/preprocess.m:20-55).  Host pointers. */
typedef struct hdgx_master {
  int32_t nd, porder, npv, npf, nfe, ngv, ngf;
  const double* shapvt;          /* (ngv, npv, nd+1) */
  const double* shapvg;          /* (npv, ngv, nd+1) */
  const double* shapvgdotshapvl; /* (npv*npv, ngv, nd+1) */
  const double* shapft;          /* (ngf, npf, nd) */
  const double* shapfg;          /* (npf, ngf, nd) */
  const double* shapfgdotshapfc; /* (npf*npf, ngf, nd) */
  const double* shapmv;          /* (ngv, npv, nd+1) geometry basis */
  const double* shapmf;          /* (ngf, npf, nd)   geometry basis */
  const int32_t* perm;           /* (npf, nfe), 0-based */
  const int32_t* permgeom;       /* (npf, nfe), 0-based */
} hdgx_master;

/* Application struct: the fields of the reference's `app` that the elemental
 * path reads (kernel/volint.m:10-23, kernel/faceint.m:13-27). */
#define HDGX_MAX_PARAM 8
#define HDGX_MAX_BC 8
typedef struct hdgx_app {
  int32_t physics;    /* HDGX_PHYS_EULER or HDGX_PHYS_NS */
  int32_t nd, ncu, nc, nch, ncd;
  int32_t localsolve; /* 0: u only (schur_u), 1: u and q (schur_uq) */
  int32_t tdep;       /* 1: add the BDF source s += SH - fc_u*u (volint.m:180-221) */
  int32_t av_col;     /* dgnodes column with the artificial viscosity, or -1 */
  int32_t getdqdg;    /* 1: nco = nc (hdg_elemental.m:20-26); only 1 is supported */
  double fc_u, fc_q, time;
  int32_t nparam;
  double param[HDGX_MAX_PARAM]; /* Euler {gam, eps, tau}; NS {gam, eps, Re, Pr, Minf, tau} */
  int32_t nbc;
  int32_t bcm[HDGX_MAX_BC];                 /* boundary model per boundary id */
  double bcs[HDGX_MAX_BC * HDGX_MAX_PARAM]; /* bcs[i*HDGX_MAX_PARAM + c] = state of id i+1 */
} hdgx_app;

/* Context: one per host thread or stream.  device = HIP device ordinal. */
int hdgx_create(hdgx_ctx** ctx, int device);
int hdgx_destroy(hdgx_ctx* ctx);
/* stream is a hipStream_t (NULL = default stream); calls are async on it */
int hdgx_set_stream(hdgx_ctx* ctx, void* stream);
int hdgx_set_precision(hdgx_ctx* ctx, int precision);
const char* hdgx_last_error(hdgx_ctx* ctx);

/* Elemental operator.  Replaces
 *   [AE,FE,DUDG,DUDG_DUH] = hdg_elemental(master,app,dgnodes,bf,UDG,UH,SH)
 * (kernel/hdg_elemental.m:1-203) called from kernel/hdg_assemble.m:41-43.
 *   dgnodes (npv, ncd, ne), bf (nfe, ne) int32 (<0: boundary id -bf),
 *   udg (npv, nc, ne), uh (ncu, npf*nfe, ne) gathered trace (hdg_assemble.m:34),
 *   sh (npv, nc, ne) (read only when app->tdep or localsolve);
 *   ae (ncu, ndf, ncu, ndf, ne), fe (ncu, ndf, ne),
 *   dudg (npv, nc, ne), dudg_duh (npv, nc, ncu, ndf, ne), status (ne) int32. */
int hdgx_elemental(hdgx_ctx* ctx, const hdgx_master* master, const hdgx_app* app, int64_t ne,
                   const double* dgnodes, const int32_t* bf, const double* udg, const double* uh,
                   const double* sh, double* ae, double* fe, double* dudg, double* dudg_duh,
                   int32_t* status);

/* Element matrices only: the outputs of volint + faceint that feed the Schur
 * step (kernel/volint.m:1-263, kernel/faceint.m:1-360), in the layouts that
 * kernel/schur_primal.m:5-11 receives them:
 *   BD (npv,ncu,npv,nc,ne), F (npv,ncu,ncu,ndf,ne), GK (ncu,ndf,npv,nc,ne),
 *   H (ncu,ndf,ncu,ndf,ne), Ru (npv,ncu,ne), Rh (ncu,ndf,ne),
 *   and for localsolve=1: M (npv,npv,ne), C (npv,npv,nd,ne), E (npv,ndf,nd,ne),
 *   Rq (npv,ncu,nd,ne) (pass NULL for these when localsolve=0). */
int hdgx_elemental_matrices(hdgx_ctx* ctx, const hdgx_master* master, const hdgx_app* app,
                            int64_t ne, const double* dgnodes, const int32_t* bf,
                            const double* udg, const double* uh, const double* sh, double* BD,
                            double* F, double* GK, double* H, double* Ru, double* Rh, double* M,
                            double* C, double* E, double* Rq);

/* u-only Schur elimination.  Replaces the MEX
 *   [DUDG,DUDG_DUH,AE,FE] = schur_u(BD,F,GK,H,Ru,Rh)   (kernel/schur_u.c:82-163)
 * with the same argument order; inputs are const (the MEX clobbers BD and Ru). */
int hdgx_schur_u(hdgx_ctx* ctx, int npv, int ncu, int nc, int ndf, int64_t ne, const double* BD,
                 const double* F, const double* GK, const double* H, const double* Ru,
                 const double* Rh, double* dudg, double* dudg_duh, double* ae, double* fe,
                 int32_t* status);

/* u,q Schur elimination.  Replaces the MEX
 *   [DUDG,DUDG_DUH,AE,FE] = schur_uq(M,C,E,BD,F,GK,H,Rq,Ru,Rh)  (kernel/schur_uq.c:193-287)
 * with the same argument order; inputs are const. */
int hdgx_schur_uq(hdgx_ctx* ctx, int nd, int npv, int ncu, int nc, int ndf, int64_t ne,
                  const double* M, const double* C, const double* E, const double* BD,
                  const double* F, const double* GK, const double* H, const double* Rq,
                  const double* Ru, const double* Rh, double* dudg, double* dudg_duh,
                  double* ae, double* fe, int32_t* status);

#ifdef __cplusplus
}
#endif

#endif /* HDGX_H */
