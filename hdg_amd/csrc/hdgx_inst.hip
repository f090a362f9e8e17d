// One (physics, nd, p) configuration of the hdgx kernels; built once per
// configuration with -DHDGX_PHYS= -DHDGX_ND= -DHDGX_P=.
#include <hip/hip_runtime.h>

#include "hdgx_launch.h"
#include "quad_kernel.hip"
#include "getq_kernel.hip"
#include "schur_mfma.hip"
#include "schur64.hip"
#include "schur_inv.hip"
#include "seg_kernels.hip"

namespace hdgx {

template <class C>
hipError_t quad_launch(hipStream_t s, const QuadArgs& a, const DevMaster& t, const DevApp& ap, bool tile) {
  if constexpr (WsLayout<C>::TILE_OK) {
    if (tile) {
      hipLaunchKernelGGL((quad_kernel<C, true>), dim3((unsigned)a.ne), dim3(QuadThreads<C>::N), 0, s, a, t, ap);
      return hipGetLastError();
    }
  }
  if (tile) return hipErrorInvalidValue;
  hipLaunchKernelGGL((quad_kernel<C, false>), dim3((unsigned)a.ne), dim3(QuadThreads<C>::N), 0, s, a, t, ap);
  return hipGetLastError();
}

template <class C>
size_t ws_words() {
  return WsLayout<C>::max_words();
}

template <class C>
hipError_t resid_launch(hipStream_t s, const QuadArgs& a, const DevMaster& t, const DevApp& ap) {
  hipLaunchKernelGGL(resid_kernel<C>, dim3((unsigned)a.ne), dim3(QuadThreads<C>::N), 0, s, a, t, ap);
  return hipGetLastError();
}

template <class C>
hipError_t getq_launch(hipStream_t s, const GetqArgs& a, const DevMaster& t) {
  hipLaunchKernelGGL(getq_kernel<C>, dim3((unsigned)a.ne), dim3(GETQ_THREADS), 0, s, a, t);
  return hipGetLastError();
}

template <class C>
hipError_t schur_launch(hipStream_t s, const SchurArgs& a, const int32_t* perm, bool tile) {
  if constexpr (InvShape<C>::USE) {
    // the library's own workspace: D^-1 replaces D, then the MFMA products
    // (face blocks with the compact layout)
    if (a.prec == 0 && a.bd_scratch) {
      constexpr bool FB = WsLayout<C>::FB;
      if (tile && !FB) return hipErrorInvalidValue;
      if (tile) hipLaunchKernelGGL((inv64_kernel<C, FB>), dim3((unsigned)a.ne), dim3(64), 0, s, a, perm);
      else hipLaunchKernelGGL((inv64_kernel<C, false>), dim3((unsigned)a.ne), dim3(64), 0, s, a, perm);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      if (tile)
        hipLaunchKernelGGL(apply64_fb_kernel<C>, dim3((unsigned)apply64_blocks<C, true>(a.ne)), dim3(64), 0, s, a, perm);
      else
        hipLaunchKernelGGL(apply64_kernel<C>, dim3((unsigned)apply64_blocks<C>(a.ne)), dim3(64), 0, s, a);
      return hipGetLastError();
    }
  }
  if constexpr (S64Shape<C>::USE) {
    if (a.prec == 0) {
      if (tile && !a.aek) return hipErrorInvalidValue;
      if (S64Shape<C>::NW > 1 && a.status)
        hipLaunchKernelGGL(s64_zero_status<C>, dim3((unsigned)((a.ne + 255) / 256)), dim3(256), 0, s, a.status, a.ne);
      const dim3 grid((unsigned)s64_blocks<C>(a.ne));
      if constexpr (WsLayout<C>::TILE_OK && !WsLayout<C>::FB) {
        if (tile) {
          hipLaunchKernelGGL((schur64_kernel<C, true>), grid, dim3(64), 0, s, a);
          return hipGetLastError();
        }
      }
      if (tile) return hipErrorInvalidValue;
      hipLaunchKernelGGL((schur64_kernel<C, false>), grid, dim3(64), 0, s, a);
      return hipGetLastError();
    }
  }
  if (tile) return hipErrorInvalidValue;
  if (a.prec == 1)
    hipLaunchKernelGGL((schur_mfma_ir_kernel<C>), dim3((unsigned)a.ne), dim3(64), 0, s, a, perm);
  else
    hipLaunchKernelGGL((schur_mfma_kernel<C>), dim3((unsigned)a.ne), dim3(64), 0, s, a, perm);
  return hipGetLastError();
}

template <class C>
hipError_t seg_launch(hipStream_t s, const SegTable& t, const DevMaster& m, const int32_t* perm) {
  if (t.total <= 0) return hipSuccess;
  hipLaunchKernelGGL(quad_seg_kernel<C>, dim3((unsigned)t.total), dim3(QuadThreads<C>::N), 0, s, t, m);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if constexpr (InvShape<C>::USE) {
    constexpr bool FB = WsLayout<C>::FB;      // the batched workspace is compact where a layout exists
    hipLaunchKernelGGL(inv_seg_kernel<C>, dim3((unsigned)t.total), dim3(64), 0, s, t, m.perm);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(apply_seg_kernel<C>, dim3((unsigned)apply64_blocks<C, FB>(t.total)), dim3(64), 0, s, t, m.perm);
  } else {
    hipLaunchKernelGGL(schur_seg_kernel<C>, dim3((unsigned)t.total), dim3(64), 0, s, t, perm, m.aek[C::UQ ? 1 : 0]);
  }
  return hipGetLastError();
}

using ThisCfg = Cfg<HDGX_PHYS, HDGX_ND, HDGX_P>;
template hipError_t quad_launch<ThisCfg>(hipStream_t, const QuadArgs&, const DevMaster&, const DevApp&, bool);
template size_t ws_words<ThisCfg>();
template hipError_t resid_launch<ThisCfg>(hipStream_t, const QuadArgs&, const DevMaster&, const DevApp&);
#if HDGX_PHYS == 0
template hipError_t getq_launch<ThisCfg>(hipStream_t, const GetqArgs&, const DevMaster&);
#endif
template hipError_t schur_launch<ThisCfg>(hipStream_t, const SchurArgs&, const int32_t*, bool);
template hipError_t seg_launch<ThisCfg>(hipStream_t, const SegTable&, const DevMaster&, const int32_t*);

}  // namespace hdgx
