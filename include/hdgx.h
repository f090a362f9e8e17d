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

/* context flags (see hdgx_set_flags) */
enum { HDGX_INPLACE = 1, HDGX_HOST_PTRS = 2 };

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
  int32_t getdqdg;    /* DUDG rows kept (hdg_elemental.m:20-26): 1 nco = nc, 0 nco = ncu (q
                         recovered by hdgx_getq), other nco = 0; batched launches: 1 only */
  double fc_u, fc_q, time;
  int32_t nparam;
  double param[HDGX_MAX_PARAM]; /* Euler {gam, eps, tau}; NS {gam, eps, Re, Pr, Minf, tau} */
  int32_t nbc;
  /* boundary model per boundary id, the reference's fbou cases: 1 far field
   * (all); 2D Euler (euler/fbou.m): 2 inviscid wall, 3 and 4 prescribed
   * pressure, 5 inflow; NS (ns/fbou.m): 2 adiabatic wall, 3 isothermal wall,
   * 4 prescribed pressure, 6 Dirichlet.  Others: HDGX_ERR_UNSUPPORTED. */
  int32_t bcm[HDGX_MAX_BC];
  double bcs[HDGX_MAX_BC * HDGX_MAX_PARAM]; /* bcs[i*HDGX_MAX_PARAM + c] = state of id i+1 */
} hdgx_app;

/* Context: one per host thread or stream.  device = HIP device ordinal. */
int hdgx_create(hdgx_ctx** ctx, int device);
int hdgx_destroy(hdgx_ctx* ctx);
/* stream is a hipStream_t (NULL = default stream); calls are async on it */
int hdgx_set_stream(hdgx_ctx* ctx, void* stream);
/* Arithmetic of the local elimination of later hdgx_elemental / hdgx_schur_*
 * calls on ctx (the reference is FP64 throughout, kernel/schur_uq.c:40-287):
 *   HDGX_PREC_FP64     FP64 Gauss-Jordan on MFMA (default);
 *   HDGX_PREC_FP32_IR  FP32 Gauss-Jordan and D'^-1 on MFMA, every solution
 *                      column refined in FP64 against the element's FP64 BD
 *                      until the predicted error is below 1e-13 relative
 *                      (HDGX_ST_IR_SLOW in status[e] if 4 sweeps do not get
 *                      there); AE, FE and the q rows stay FP64.
 * The quadrature is FP64 in both. */
int hdgx_set_precision(hdgx_ctx* ctx, int precision);
/* Flags of later hdgx_schur_u / hdgx_schur_uq calls on ctx (0 = none):
 *   HDGX_INPLACE  also overwrite the inputs the way the MEX files do, for
 *                 comparisons with them (kernel/schur_uq.c:69-136,
 *                 kernel/schur_u.c:67): M -> its dpotrf('U') factor (upper
 *                 triangle; the strict lower one untouched), C, E, Rq ->
 *                 M^-1 C, M^-1 E, M^-1 Rq, Ru -> Ru - B M^-1 Rq, F -> B M^-1 E
 *                 - F, and the D block of BD -> dgetrf's LU of D + B M^-1 C
 *                 (schur_u: of D = BD).  The pointers are declared const
 *                 but are written.  One more launch per call; the outputs
 *                 are unchanged.  hdgx_elemental is unaffected (its Schur
 *                 inputs are internal).
 *   HDGX_HOST_PTRS  the array arguments of hdgx_elemental,
 *                 hdgx_elemental_matrices, hdgx_residual, hdgx_getq,
 *                 hdgx_schur_u and hdgx_schur_uq are host memory: they are
 *                 copied to a device staging buffer owned by ctx, the call
 *                 runs there, and the outputs (and the arrays HDGX_INPLACE
 *                 writes) are copied back before it returns.  Without the
 *                 flag every array is device memory and calls are
 *                 asynchronous.  The trace-solve entry points always take
 *                 device memory. */
int hdgx_set_flags(hdgx_ctx* ctx, int flags);
/* Most device bytes of the context's element-matrix workspace, which carries
 * the element matrices from the quadrature to the elimination inside
 * hdgx_elemental / hdgx_elemental_batch (0 = the default, 24 GiB).  A call
 * whose elements need more runs in chunks of nearly equal size that fit it,
 * as the reference blocks its element loop (kernel/hdg_elemental.m:9-13);
 * results do not depend on the chunking.  The workspace is allocated at the
 * first call that needs it, grows up to the limit, and is held until
 * hdgx_destroy (or freed here when it is above a new, lower limit). */
int hdgx_set_workspace_limit(hdgx_ctx* ctx, int64_t bytes);
/* Kernel timing of later hdgx_elemental / hdgx_elemental_noalloc calls on ctx
 * (no counterpart in the reference; for measurement): with on != 0 each
 * chunk's quadrature and elimination launches are bracketed by HIP events on
 * the context's stream and the call synchronises to add their elapsed times;
 * hdgx_kernel_times returns the sums (milliseconds) and the number of
 * launch pairs since the last hdgx_set_timing. */
int hdgx_set_timing(hdgx_ctx* ctx, int on);
int hdgx_kernel_times(hdgx_ctx* ctx, double* quad_ms, double* elim_ms, int64_t* launches);
const char* hdgx_last_error(hdgx_ctx* ctx);

/* Elemental operator.  Replaces
 *   [AE,FE,DUDG,DUDG_DUH] = hdg_elemental(master,app,dgnodes,bf,UDG,UH,SH)
 * (kernel/hdg_elemental.m:1-203) called from kernel/hdg_assemble.m:41-43.
 *   dgnodes (npv, ncd, ne), bf (nfe, ne) int32 (<0: boundary id -bf),
 *   udg (npv, nc, ne), uh (ncu, npf*nfe, ne) gathered trace (hdg_assemble.m:34),
 *   sh (npv, nc, ne) (read only when app->tdep or localsolve);
 *   ae (ncu, ndf, ncu, ndf, ne), fe (ncu, ndf, ne),
 *   dudg (npv, nco, ne), dudg_duh (npv, nco, ncu, ndf, ne) with nco = nc, ncu or 0
 *   as app->getdqdg is 1, 0 or other (may be NULL when nco = 0), status (ne) int32. */
int hdgx_elemental(hdgx_ctx* ctx, const hdgx_master* master, const hdgx_app* app, int64_t ne,
                   const double* dgnodes, const int32_t* bf, const double* udg, const double* uh,
                   const double* sh, double* ae, double* fe, double* dudg, double* dudg_duh,
                   int32_t* status);

/* Elemental operator that allocates nothing.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory -- code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Arguments, layouts, results and kernels are hdgx_elemental's
 * (bit for bit); every array is device memory (HDGX_HOST_PTRS is refused).
 * The master tables and the element matrices between the two kernels live in
 * a 128 MiB device array of the library (1 MiB of tables, the rest
 * workspace), which every GPU the library runs on holds until the process
 * ends; the elements run in chunks that fit it, so ne is not limited, but a
 * master whose tables pass 1 MiB is refused (HDGX_ERR_ARG).  The call makes
 * no hipMalloc / hipFree.  When its master differs from the last one used on
 * that GPU it first copies the tables in, which synchronises ctx's stream.
 * Every call on one GPU shares the array: calls that overlap on one GPU
 * (other streams, threads or contexts) race and must be serialised. */
int hdgx_elemental_noalloc(hdgx_ctx* ctx, const hdgx_master* master, const hdgx_app* app, int64_t ne,
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

/* Residual only, no Jacobians.  Replaces the element part of
 *   [RU,RH,RQ] = hdg_residual(master,app,mesh,UDG,UH,SH)  (kernel/hdg_residual.m:1-37)
 * with its volint / faceint (hdg_residual.m:57-290): Ru (npv,ncu,ne) and Rq
 * (npv,ncu,nd,ne; NULL when localsolve=0) as hdgx_elemental_matrices gives them,
 * Rh (ncu,ndf,ne) per element with hdg_residual's sign, +shapfg (fbou | fhat)
 * (hdg_residual.m:286-290; kernel/faceint.m:339 negates it).  The global RH of
 * hdg_residual.m:41-52 is hdgx_trace_scatter of this Rh.  Inputs as
 * hdgx_elemental; uh is the element-gathered trace (hdgx_trace_gather). */
int hdgx_residual(hdgx_ctx* ctx, const hdgx_master* master, const hdgx_app* app, int64_t ne,
                  const double* dgnodes, const int32_t* bf, const double* udg, const double* uh,
                  const double* sh, double* Ru, double* Rh, double* Rq);

/* q recovery.  Replaces
 *   [QDG, qq] = getq(master, mesh, UDG, UH, SH, fc_q)     (kernel/getq.m:1-129)
 * as kernel/hdg_solve.m:130,149 and kernel/hdg_solve_dirk.m:89 call it:
 *   QDG = (fc_q M)^-1 (C u - E uh) + SH(:, end-ncu*nd+1:end, :)/fc_q
 * per element, the q that zeroes Rq of volint.m:57-60 / faceint.m:48-51.
 *   dgnodes (npv,ncd,ne); udg (npv,nc,ne), columns 0..ncu-1 read; uh the
 *   element-gathered trace (ncu,npf*nfe,ne) (hdgx_trace_gather); sh
 *   (npv,nc,ne) or NULL.  Element e's (npv, ncu*nd) block of QDG is written
 *   at qdg + e*ldq: ldq = npv*ncu*nd for a separate array, or qdg = udg +
 *   npv*ncu with ldq = npv*nc to overwrite UDG(:, ncu+1:end, :) in place
 *   (hdg_solve.m:131).  shapnv (npv,npv,nd+1) = mkshape(porder, plocvl,
 *   plocvl) and qq (npv,ncu*nd,ne), the nodal gradient -grad u of
 *   getq.m:62-76, are both given or both NULL.  fc_q must be nonzero.
 * getq.m's MiCE cache (Mi, C, E per element) has no counterpart: the kernel
 * rebuilds the geometry, which costs less than reading those matrices. */
int hdgx_getq(hdgx_ctx* ctx, const hdgx_master* master, int ncu, int nc, int ncd, int64_t ne,
              const double* dgnodes, const double* udg, const double* uh, const double* sh,
              double fc_q, double* qdg, int64_t ldq, const double* shapnv, double* qq);

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

/* One request's element range in a batched elemental launch.  Device arrays
 * laid out as for hdgx_elemental; app is a host pointer. */
typedef struct hdgx_segment {
  int64_t ne;
  const double* dgnodes;
  const int32_t* bf;
  const double* udg;
  const double* uh;
  const double* sh;
  double* ae;
  double* fe;
  double* dudg;
  double* dudg_duh;
  int32_t* status;
  const hdgx_app* app;
} hdgx_segment;

/* Elemental operator over the element ranges of many requests in one batched
 * launch per kernel (FP64): each segment is the hdg_elemental call of one
 * request (kernel/hdg_elemental.m:1-203, as called per mesh partition by
 * kernel/hdg_assemble.m:40-43), with its own physics parameters.  All segments
 * share the master (nd, p) and the app's physics, nd, ncu, nc and localsolve;
 * results equal those of hdgx_elemental on each segment alone (bit for bit in
 * 2D; for tetrahedra to rounding, 1e-13 relative, as the batched quadrature
 * kernel contracts a few multiply-adds differently). */
int hdgx_elemental_batch(hdgx_ctx* ctx, const hdgx_master* master, int nseg,
                         const hdgx_segment* segs);

/* The same over segments of several polynomial orders (one dimension):
 * masters[s] is the master of segment s.  Segments are grouped by order and
 * each group is one hdgx_elemental_batch (one batched launch per kernel),
 * the groups back to back on the context's stream, highest order first;
 * results equal those of hdgx_elemental_batch on each group.  For callers
 * whose batches change every call (a serving loop) -- a fixed batch replayed
 * many times is a launch plan (below).  No counterpart in the reference, whose
 * element loop runs one mesh at a time (kernel/hdg_assemble.m:40-43). */
int hdgx_elemental_batch_mixed(hdgx_ctx* ctx, int nseg, const hdgx_master* const* masters,
                               const hdgx_segment* segs);

/* Launch plan: a batch of segments of several polynomial orders (one physics
 * and dimension), with tables and element-matrix workspace set up once so
 * that the same batch can be launched many times, and captured into a
 * hipGraph.  masters[s] is the master of segment s (its order p).  Modes:
 *   HDGX_PLAN_PER_ORDER   one batched launch per order, in sequence;
 *   HDGX_PLAN_CONCURRENT  one batched launch per order, on concurrent streams;
 *   HDGX_PLAN_FUSED       one persistent launch per kernel over all orders
 *                         (grid workgroups; 0 = 4 per CU).
 * Results are those of hdgx_elemental on each segment (FP64): bit for bit in
 * the per-order and concurrent modes; the fused mode eliminates every order
 * in the FP64 MFMA form, so where hdgx_elemental takes the LDS form
 * (tetrahedra p = 2) they agree to rounding (1e-12 relative).  The plan
 * owns its workspace; it has no counterpart in the reference, whose element
 * loop runs one mesh at a time (kernel/hdg_assemble.m:40-43). */
enum { HDGX_PLAN_PER_ORDER = 0, HDGX_PLAN_FUSED = 1, HDGX_PLAN_CONCURRENT = 2 };
typedef struct hdgx_plan hdgx_plan;
int hdgx_plan_create(hdgx_ctx* ctx, int nseg, const hdgx_master* const* masters,
                     const hdgx_segment* segs, int mode, int grid, hdgx_plan** plan);
/* enqueue the batch on the context's stream (replays the graph once captured) */
int hdgx_plan_launch(hdgx_ctx* ctx, hdgx_plan* plan);
/* capture the plan's launches into a hipGraph used by later hdgx_plan_launch */
int hdgx_plan_capture(hdgx_ctx* ctx, hdgx_plan* plan);
int hdgx_plan_destroy(hdgx_ctx* ctx, hdgx_plan* plan);

/* ------------------------------------------------------------------------
 * Trace system and Newton update (kernel/hdg_solve.m:105-176 with
 * app.denseblock = 1).  The global trace unknown UH is (ncu, npf, nf): face i
 * owns trace nodes i*npf .. i*npf+npf-1 (elcon numbering).  The system is kept
 * as face-block rows: row i holds the (ncf x ncf) blocks, ncf = ncu*npf, of the
 * faces that share an element with face i.  Slot 0 is face i itself, slots
 * 1..nbf-1 are the other faces of its first element, then of its second
 * element, in ascending local-face order (mesh/mkmesh/mkf2f.m:1-60),
 * nbf = 2*nfe - 1.
 * ------------------------------------------------------------------------ */
typedef struct hdgx_trace {
  int32_t ncu, npf, nfe;
  int64_t ne, nf;
  const int32_t* fconn;      /* (4, nf) device: e1, j1, e2, j2 per face (0-based; e2 = -1 on the boundary) */
  const int32_t* f2f;        /* (nbf, nf) device: face of each block slot, -1 where none */
  const int32_t* elcon;      /* (npf*nfe, ne) device: element trace node -> global trace node */
  /* multicolour ordering of the face graph, used by the block ILU(0) */
  int32_t ncolors;
  const int32_t* color_ptr;  /* (ncolors+1) HOST: faces of colour c are color_rows[color_ptr[c] .. color_ptr[c+1]) */
  const int32_t* color_rows; /* (nf) device */
  const int32_t* face_rank;  /* (nf) device: position of each face in the elimination order */
} hdgx_trace;

/* Face-block system from the element matrices.  Replaces
 *   [K,F] = mkDenseBlockSystem(AE, FE, mesh.f, mesh.t2f, mesh.elcon)
 * (kernel/mkDenseBlockSystem.m:13-95, called by kernel/hdg_assemble.m:78-80):
 * each block row is scaled by the inverse of its diagonal block.
 *   AE (ncu,ndf,ncu,ndf,ne), FE (ncu,ndf,ne) ->
 *   K (ncf, ncf, nbf-1, nf) (= the reference's (ncf, ncf*(nbf-1), nf)), F (ncf, nf),
 *   status (nf) int32: HDGX_ST_SINGULAR_D where the diagonal block is singular. */
int hdgx_trace_assemble(hdgx_ctx* ctx, const hdgx_trace* T, const double* AE, const double* FE,
                        double* K, double* F, int32_t* status);

/* y = x + sum_k K_k x(f2f(k)) -- the operator of kernel/gmres.m:38-42 and its
 * per-face block GEMV kernel/arraymul.c:53-70.  x, y (ncf, nf). */
int hdgx_trace_matvec(hdgx_ctx* ctx, const hdgx_trace* T, const double* K, const double* x,
                      double* y);

/* Block ILU(0) of the face-block system (identity diagonal) in the
 * multicolour elimination order of T.  Replaces kernel/computeBILU0.m:1-45.
 *   LU (ncf, ncf, nbf, nf): slot 0 the inverse of U_ii, slot k the L_ik
 *   (earlier face) or U_ik (later face) block.  status (nf): singular U_ii. */
int hdgx_bilu0_factor(hdgx_ctx* ctx, const hdgx_trace* T, const double* K, double* LU,
                      int32_t* status);

/* y = (LU)^-1 x (kernel/applyBILU0.m:1-55).  x, y (ncf, nf), may not alias. */
int hdgx_bilu0_apply(hdgx_ctx* ctx, const hdgx_trace* T, const double* LU, const double* x,
                     double* y);

/* GMRES(restart) on the face-block system, FP64, classical Gram-Schmidt
 * (kernel/gmres.m:1-111); with LU != NULL left-preconditioned by the block
 * ILU(0) (kernel/gmres_bilu0.m:1-103).  x (ncf, nf) holds the initial guess
 * and receives the solution.  Synchronous: returns when converged.
 * iters / relres / flag (0 converged, 1 maxit) may be NULL. */
int hdgx_gmres(hdgx_ctx* ctx, const hdgx_trace* T, const double* K, const double* LU,
               const double* F, double* x, int restart, double tol, int maxit, int* iters,
               double* relres, int* flag);

/* UHE(:, :, e) = UH(:, elcon(:, e)): the element-gathered trace of
 * kernel/hdg_assemble.m:34 and kernel/hdg_solve.m:120.
 *   UH (ncu, nsiz), UHE (ncu, npf*nfe, ne). */
int hdgx_trace_gather(hdgx_ctx* ctx, int ncu, int npf, int nfe, int64_t ne, const int32_t* elcon,
                      const double* UH, double* UHE);

/* Greedy colouring of the face graph for the block-ILU(0) row order
 * (hdgx_trace.color_ptr / color_rows): face i, in order, takes the smallest
 * colour none of its neighbours f2f(1:nbf-1, i) (0-based, -1 none) already
 * has.  Host memory; no context.  The reference orders the rows by minimum
 * discarded fill (kernel/computeMDFordering.m), which is sequential; a
 * colour order lets each colour's rows be factored and solved in parallel.
 *   f2f (nbf, nf) int32, color (nf) int32 out, *ncolors out. */
int hdgx_color_faces(int nbf, int64_t nf, const int32_t* f2f, int32_t* color, int32_t* ncolors);

/* Mesh faces on the device (mesh/mkmesh/mkt2f.m): faces numbered in order of
 * first visit over (element e, local face j), 0-based.
 *   t (nd+1, ne) int32 vertex ids 0..nvert-1 (3D: nvert <= 2^21);
 *   t2f (nd+1, ne) int32 out: face of local face j of element e;
 *   f (nd+2, ne*(nd+1)) int32 out, column i = face i for i < *nf: its nd
 *     vertices as its first element lists them (the local face tables of
 *     mkt2f.m), the first element, the second element or -1 on the boundary;
 *   *nf out (host) the number of faces.
 * The reference's f and t2f are 1-based and f's last column later holds -bnd
 * ids (setbndnbrs); these are 0-based and keep -1.  Device arrays; the call
 * synchronises the context's stream.  HDGX_ERR_ARG for a vertex id out of
 * range or a face shared by more than two elements. */
int hdgx_mesh_faces(hdgx_ctx* ctx, int nd, int64_t ne, int32_t nvert, const int32_t* t, int32_t* t2f, int32_t* f,
                    int64_t* nf);

/* Trace-node numbering on the device (mesh/mkmesh/elconnectivities.m:147-212):
 * face i owns trace nodes i*npf .. i*npf+npf-1 in its first element's order;
 * each node of the second element's face takes the first-side node at the
 * least squared distance over the first min(ncd, 3) coordinates (an error
 * past 1e-16).
 *   dgnodes (npv, ncd, ne) FP64; perm (npf, nd+1) int32 (master.perm,
 *   0-based); f, t2f as hdgx_mesh_faces gives them; elcon (npf*(nd+1), ne)
 *   int32 out.  Device arrays; synchronises the context's stream.  An element
 *   id of f outside 0..ne-1 or a perm entry outside 0..npv-1 is HDGX_ERR_ARG. */
int hdgx_mesh_elcon(hdgx_ctx* ctx, int nd, int npv, int ncd, int npf, int64_t ne, int64_t nf, const double* dgnodes,
                    const int32_t* perm, const int32_t* f, const int32_t* t2f, int32_t* elcon);

/* Face-to-face table on the device (mesh/mkmesh/mkf2f.m:8-55, 0-based):
 * f2f (2nd+1, nf) int32 out: the face, the other faces of its first element,
 * then of its second, in ascending local index, -1 where there is no second
 * element; fconn (4, nf) int32 out (may be NULL): first element, its local
 * face, second element or -1, its local face or 0 (hdgx_trace.fconn).
 * f, t2f as hdgx_mesh_faces gives them for ne elements (a negative second
 * element, e.g. a boundary tag, means none; an element id of f at or past ne
 * is HDGX_ERR_ARG).  Device arrays; synchronises the stream. */
int hdgx_mesh_f2f(hdgx_ctx* ctx, int nd, int64_t ne, int64_t nf, const int32_t* f, const int32_t* t2f, int32_t* fconn,
                  int32_t* f2f);

/* CBSR pattern of the face-block trace system on the device
 * (mesh/mkmesh/block_crs.m:45-62 with elcon2entcon.m on t2f', 0-based): row i
 * is face i, then the other faces sharing an element with it, ascending.
 *   f2f (2nd+1, nf) int32 as hdgx_mesh_f2f gives it; rowpts (nf+1) int32 out,
 *   offsets from 0 (cbsr_rowpts); colind int32 out (cbsr_colind, 0-based face
 *   ids), room for nf*(2nd+1) entries; *nblks = rowpts(nf) (cbsr_nblks).
 * Device arrays; synchronises the context's stream.  HDGX_ERR_ARG when
 * f2f(0, i) is not i or another entry is i or at least nf (colind is then not
 * written). */
int hdgx_mesh_block_crs(hdgx_ctx* ctx, int nd, int64_t nf, const int32_t* f2f, int32_t* rowpts, int32_t* colind,
                        int64_t* nblks);

/* RH(:, elcon(:, e)) += Rh(:, :, e) over all elements: the global trace
 * residual of kernel/hdg_residual.m:41-52 (sparse(il, 1, RH)), summed per face
 * from its one or two elements (no atomics, the same bits every run).
 *   Rh (ncu, npf*nfe, ne), RH (ncu, npf*nf) in the layout of UH. */
int hdgx_trace_scatter(hdgx_ctx* ctx, const hdgx_trace* T, const double* Rh, double* RH);

/* Element-by-element local solves: DUDGT = dudg + dudg_duh * DUH(:, elcon)
 * recomputed per element from the state, without DUDG_DUH for the whole mesh
 * (replaces kernel/hdg_local.m:1-85, hdg_solve.m:122-124, the path of
 * app.getdqdg outside {0, 1}).  Per workspace-sized chunk of elements it runs
 * the elemental kernels (volint, faceint and the elimination: schur_primal's
 * dudg, dudg_duh) into scratch and contracts them with the chunk's trace
 * increments, as hdgx_local_update does.  Inputs as hdgx_elemental (uh the
 * gathered trace (ncu, npf*nfe, ne)), elcon (npf*nfe, ne) 0-based, DUH
 * (ncu, nsiz); DUDGT (npv, nc, ne) keeps every row; status (ne) optional.
 * Device arrays only. */
int hdgx_local_solve(hdgx_ctx* ctx, const hdgx_master* master, const hdgx_app* app, int64_t ne,
                     const double* dgnodes, const int32_t* bf, const double* udg, const double* uh, const double* sh,
                     const int32_t* elcon, const double* DUH, double* dudgt, int32_t* status);

/* Local update DUDGT = DUDG + DUDG_DUH * DUH(:, elcon) (kernel/hdg_solve.m:119-121,
 * util/mapContractK.m:47-99).  DUDG (npv,nco,ne), DUDG_DUH (npv,nco,ncu,ndf,ne),
 * DUH (ncu, nsiz), DUDGT (npv,nco,ne). */
int hdgx_local_update(hdgx_ctx* ctx, int npv, int nco, int ncu, int npf, int nfe, int64_t ne,
                      const int32_t* elcon, const double* DUDG, const double* DUDG_DUH,
                      const double* DUH, double* DUDGT);

/* Damped Newton step (kernel/hdg_solve.m:145-146):
 *   UDG(:, 1:nco, :) = UDG0(:, 1:nco, :) + alpha*DUDGT,  UH = UH0 + alpha*DUH.
 * UDG, UDG0 (npv, nc, ne); DUDGT (npv, nco, ne); UH, UH0, DUH (nuh).  UDG may
 * alias UDG0 and UH may alias UH0. */
int hdgx_newton_update(hdgx_ctx* ctx, int npv, int nc, int nco, int64_t ne, int64_t nuh,
                       double alpha, const double* UDG0, const double* DUDGT, const double* UH0,
                       const double* DUH, double* UDG, double* UH);

/* Euclidean norm of a device vector, synchronous (norm(F(:)) of
 * kernel/hdg_solve.m:105,154). */
int hdgx_norm2(hdgx_ctx* ctx, int64_t n, const double* x, double* result);

#ifdef __cplusplus
}
#endif

#endif /* HDGX_H */
