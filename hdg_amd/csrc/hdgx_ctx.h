// Context object of the hdgx C ABI and the error helpers shared by the API
// translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>

#include "../../include/hdgx.h"
#include "hdgx_kernels.h"

struct MasterBuf {
  void* dev = nullptr;
  hdgx::DevMaster tab{};
  int nd = 0, porder = 0, npv = 0, npf = 0, nfe = 0, ngv = 0, ngf = 0;
};

struct hdgx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int prec = HDGX_PREC_FP64;
  int flags = 0;                 // HDGX_INPLACE, HDGX_HOST_PTRS
  std::string err;
  std::map<uint64_t, MasterBuf> masters;
  void* ws = nullptr;            // element matrices between the quadrature and Schur kernels
  size_t ws_bytes = 0;
  size_t ws_cap = (size_t)24 << 30;  // most bytes ws may hold (hdgx_set_workspace_limit)
  void* tws = nullptr;           // trace-solve workspace (Krylov basis, reductions)
  size_t tws_bytes = 0;
  double* pinned = nullptr;      // small host staging for reductions
  void* segdev = nullptr;        // segment and physics tables of batched launches
  size_t segdev_bytes = 0;
  void* hws = nullptr;           // device copies of host arrays (HDGX_HOST_PTRS)
  size_t hws_bytes = 0;
  void* mws = nullptr;           // mesh-connectivity scratch (hash table, scan)
  size_t mws_bytes = 0;
  // kernel timing of hdgx_elemental (hdgx_set_timing): events around the
  // quadrature and elimination launches, summed milliseconds
  bool timing = false;
  hipEvent_t tev[3] = {nullptr, nullptr, nullptr};
  double tsum[2] = {0.0, 0.0};
  int64_t tcalls = 0;
};

namespace hdgx_detail {

inline int fail(hdgx_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

}  // namespace hdgx_detail

#define HIP_TRY(ctx, call)                                                       \
  do {                                                                           \
    hipError_t _e = (call);                                                      \
    if (_e != hipSuccess)                                                        \
      return hdgx_detail::fail(ctx, HDGX_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)
