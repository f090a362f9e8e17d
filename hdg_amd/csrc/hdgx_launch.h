// Per-configuration launchers.  Each configuration is compiled in its own
// translation unit (hdgx_inst.hip with HDGX_PHYS/HDGX_ND/HDGX_P defined), so
// the kernels build in parallel; the API layer only sees these declarations.
#pragma once

#include <hip/hip_runtime.h>

#include "hdgx_kernels.h"

// The configurations built into the library: (physics, nd, p).
#define HDGX_FOR_EACH_CFG(X) \
  X(0, 2, 1) X(0, 2, 2) X(0, 2, 3) X(0, 2, 4) \
  X(1, 2, 1) X(1, 2, 2) X(1, 2, 3) X(1, 2, 4) \
  X(0, 3, 1) X(0, 3, 2)

// The (physics, nd) families with a mixed-order persistent launch.
#define HDGX_FOR_EACH_FAMILY(X) X(0, 2) X(1, 2) X(0, 3)

namespace hdgx {

// tile: write F, GK and H in the tile layout (WsLayout in hdgx_shapes.h) for a
// schur_launch with tile set; otherwise the reference layout
template <class C> hipError_t quad_launch(hipStream_t s, const QuadArgs& a, const DevMaster& t, const DevApp& ap,
                                          bool tile);
// residual only (Ru, Rh with hdg_residual's sign, Rq)
template <class C> hipError_t resid_launch(hipStream_t s, const QuadArgs& a, const DevMaster& t, const DevApp& ap);
// the MEX files' in-place side effects (one launch after the elimination)
hipError_t inplace_launch(hipStream_t s, const InplaceArgs& a);
// q recovery; geometry only, so it is built for the Euler configurations
// (PH = 0), whose (nd, p) cover every built order
template <class C> hipError_t getq_launch(hipStream_t s, const GetqArgs& a, const DevMaster& t);
// prec 0: FP64 MFMA Gauss-Jordan; 1: FP32 MFMA Gauss-Jordan (with D'^-1) + FP64
// refinement.
// perm (device, may be null) marks GK as face sparse; tile: F, GK and H in the
// tile layout (FP64, configurations with WsLayout<C>::TILE_OK)
template <class C> hipError_t schur_launch(hipStream_t s, const SchurArgs& a, const int32_t* perm, bool tile);
// per-element words of the element-matrix workspace (the larger of its two layouts)
template <class C> size_t ws_words();
// batched launches over a segment table (FP64 MFMA elimination)
template <class C> hipError_t seg_launch(hipStream_t s, const SegTable& t, const DevMaster& m, const int32_t* perm);

#define HDGX_DECLARE(PH, ND, P)                                                                          \
  extern template hipError_t quad_launch<Cfg<PH, ND, P>>(hipStream_t, const QuadArgs&, const DevMaster&, \
                                                         const DevApp&, bool);                            \
  extern template size_t ws_words<Cfg<PH, ND, P>>();                                                      \
  extern template hipError_t resid_launch<Cfg<PH, ND, P>>(hipStream_t, const QuadArgs&, const DevMaster&, \
                                                          const DevApp&);                                 \
  extern template hipError_t schur_launch<Cfg<PH, ND, P>>(hipStream_t, const SchurArgs&, const int32_t*, bool); \
  extern template hipError_t seg_launch<Cfg<PH, ND, P>>(hipStream_t, const SegTable&, const DevMaster&,  \
                                                        const int32_t*);
HDGX_FOR_EACH_CFG(HDGX_DECLARE)
#undef HDGX_DECLARE
#define HDGX_DECLARE_GETQ(PH, ND, P) \
  extern template hipError_t getq_launch<Cfg<PH, ND, P>>(hipStream_t, const GetqArgs&, const DevMaster&);
HDGX_FOR_EACH_CFG(HDGX_DECLARE_GETQ)
#undef HDGX_DECLARE_GETQ

// one persistent launch per kernel over the elements of every order
template <int PH, int ND> hipError_t mixed_launch(hipStream_t s, const MixedTable* t, int64_t total, int grid);
#define HDGX_DECLARE_FAMILY(PH, ND) \
  extern template hipError_t mixed_launch<PH, ND>(hipStream_t, const MixedTable*, int64_t, int);
HDGX_FOR_EACH_FAMILY(HDGX_DECLARE_FAMILY)
#undef HDGX_DECLARE_FAMILY

}  // namespace hdgx
