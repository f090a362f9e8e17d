// One (physics, nd) family of the mixed-order persistent launches; built once
// per family with -DHDGX_PHYS= -DHDGX_ND=.
#include <hip/hip_runtime.h>

#include "hdgx_launch.h"
#include "quad_kernel.hip"
#include "schur_mfma.hip"
#include "schur64.hip"
#include "schur_inv.hip"
#include "seg_kernels.hip"
#include "mixed_kernels.hip"

namespace hdgx {

// t: the table in device memory; total: its element count
template <int PH, int ND>
hipError_t mixed_launch(hipStream_t s, const MixedTable* t, int64_t total, int grid) {
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL((quad_mixed_kernel<PH, ND>), dim3(grid), dim3(QUAD_THREADS), 0, s, t);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((schur_mixed_kernel<PH, ND>), dim3(grid), dim3(64), 0, s, t);
  return hipGetLastError();
}

template hipError_t mixed_launch<HDGX_PHYS, HDGX_ND>(hipStream_t, const MixedTable*, int64_t, int);

}  // namespace hdgx
