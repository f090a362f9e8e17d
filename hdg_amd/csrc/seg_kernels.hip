// Batched launches over the element ranges of many requests: one workgroup
// per batch element finds its request by binary search in the segment table
// and runs the same per-element bodies as the single-batch kernels, with its
// own physics parameters.
#include <hip/hip_runtime.h>

#include "hdgx_device.h"
#include "hdgx_kernels.h"
#include "hdgx_shapes.h"

namespace hdgx {

// wavefront-uniform loads of the tables through the scalar cache
template <class T>
HDGX_DEV T uload_obj(const T* p) {
  static_assert(sizeof(T) % 8 == 0, "tables are loaded in 8-byte words");
  T r;
  const __attribute__((address_space(4))) uint64_t* q = (const __attribute__((address_space(4))) uint64_t*)p;
  uint64_t* d = reinterpret_cast<uint64_t*>(&r);
#pragma unroll
  for (unsigned i = 0; i < sizeof(T) / 8; ++i) d[i] = q[i];
  return r;
}

HDGX_DEV int find_seg(const ElemSeg* segs, int nseg, int64_t g) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t e0 = *((const __attribute__((address_space(4))) int64_t*)&segs[mid].e0);
    if (e0 <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <class C>
__global__ __launch_bounds__(QuadThreads<C>::N) __attribute__((amdgpu_waves_per_eu(QuadWpe<C>::N))) void quad_seg_kernel(SegTable S, DevMaster T) {
  constexpr int NPV = C::NPV, NFE = C::NFE, NC = C::NC, NU = C::NU, NH = C::NH, NCOL = C::NCOL;
  constexpr int ND = C::ND, NDF = C::NDF, NCU = C::NCU;
  const int64_t w = blockIdx.x;
  if (w >= S.total) return;
  const int64_t g = S.g0 + w;
  const ElemSeg sg = uload_obj(S.segs + find_seg(S.segs, S.nseg, g));
  const DevApp& ap = S.apps[sg.app];
  const int64_t e = g - sg.e0;
  QuadArgs A;
  A.ne = 1;
  A.dgnodes = sg.dgnodes + e * NPV * ap.ncd;
  A.bf = sg.bf + e * NFE;
  A.udg = sg.udg + e * NPV * NC;
  A.uh = sg.uh + e * NH;
  A.sh = sg.sh ? sg.sh + e * NPV * NC : nullptr;
  // the workspace of a batched launch has the tile layout where the
  // elimination takes it (WsLayout<C>::TILE_OK)
  constexpr bool TILE = WsLayout<C>::TILE_OK;
  using L = WsLayout<C>;
  A.BD = S.BD + w * L::bd;
  A.F = S.F + w * L::f(TILE);
  A.GK = S.GK + w * L::gk(TILE);
  A.H = S.H + w * L::h(TILE);
  A.Ru = S.Ru + w * NU;
  A.Rh = S.Rh + w * NH;
  if constexpr (C::UQ) {
    A.M = S.M + w * NPV * NPV;
    A.C = S.C + w * NPV * NPV * ND;
    A.E = S.E + w * NPV * NDF * ND;
    A.Rq = S.Rq + w * NPV * NCU * ND;
  } else {
    A.M = A.C = A.E = A.Rq = nullptr;
  }
  A.status = nullptr;
  __shared__ QuadLds<C> lds;
  quad_body<C, false, TILE, QuadThreads<C>::N>(A, T, ap, 0, lds);
}

// the elimination arguments of batch element g (workspace slot w) as a
// one-element SchurArgs
template <class C>
HDGX_DEV SchurArgs seg_schur_args(const SegTable& S, int64_t w) {
  constexpr int NPV = C::NPV, NU = C::NU, NH = C::NH, NCOL = C::NCOL, ND = C::ND, NDF = C::NDF, NCU = C::NCU;
  const int64_t g = S.g0 + w;
  const ElemSeg sg = uload_obj(S.segs + find_seg(S.segs, S.nseg, g));
  const int64_t e = g - sg.e0;
  SchurArgs A;
  A.ne = 1;
  constexpr bool TILE = WsLayout<C>::TILE_OK;
  using L = WsLayout<C>;
  A.BD = S.BD + w * L::bd;
  A.F = S.F + w * L::f(TILE);
  A.GK = S.GK + w * L::gk(TILE);
  A.H = S.H + w * L::h(TILE);
  A.Ru = S.Ru + w * NU;
  A.Rh = S.Rh + w * NH;
  if constexpr (C::UQ) {
    A.M = S.M + w * NPV * NPV;
    A.C = S.C + w * NPV * NPV * ND;
    A.E = S.E + w * NPV * NDF * ND;
    A.Rq = S.Rq + w * NPV * NCU * ND;
  } else {
    A.M = A.C = A.E = A.Rq = nullptr;
  }
  A.dudg = sg.dudg + e * NCOL;
  A.dudg_duh = sg.dudg_duh + e * NCOL * NH;
  A.ae = sg.ae + e * NH * NH;
  A.fe = sg.fe + e * NH;
  A.status = sg.status ? sg.status + e : nullptr;
  A.prec = 0;
  A.nco = C::NC;
  A.aek = nullptr;
  A.bd_scratch = 1;                // the launch's workspace
  return A;
}

// the elimination body of hdgx_elemental's FP64 path for this configuration
template <class C>
constexpr int seg_wpe() { return S64Shape<C>::USE ? S64Shape<C>::WPE_SEG : MfmaShape<C>::WPE; }

template <class C>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(seg_wpe<C>(), seg_wpe<C>()))) void schur_seg_kernel(SegTable S, const int32_t* __restrict__ perm, const unsigned char* __restrict__ aek) {
  // (the four-row-tile u-only configurations take inv_seg_kernel and
  // apply_seg_kernel instead, so schur64 runs here one wavefront an element)
  const int64_t w = blockIdx.x;
  if (w >= S.total) return;
  SchurArgs A = seg_schur_args<C>(S, w);
  if constexpr (S64Shape<C>::USE && S64Shape<C>::NW == 1) {
    A.aek = aek;
    __shared__ S64Lds<C> lds;
    schur64_body<C, WsLayout<C>::TILE_OK>(A, 0, 0, lds);
  } else {
    __shared__ SchurLds<C> lds;
    schur_mfma_body<C>(A, perm, 0, lds);
  }
}

// the four-row-tile u-only configurations: D^-1 over the workspace's D, then
// the MFMA products (schur_inv.hip), as hdgx_elemental does
template <class C>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3))) void inv_seg_kernel(SegTable S,
                                                                                               const int32_t* perm) {
  if constexpr (InvShape<C>::USE) {
    const int64_t w = blockIdx.x;
    if (w >= S.total) return;
    const SchurArgs A = seg_schur_args<C>(S, w);
    __shared__ InvLds<C> lds;
    inv64_body<C, WsLayout<C>::FB>(A, 0, perm, lds);
  }
}

template <class C>
__global__ __launch_bounds__(64) void apply_seg_kernel(SegTable S, const int32_t* perm) {
  if constexpr (InvShape<C>::USE) {
    constexpr bool FB = WsLayout<C>::FB;
    constexpr int NW = FB ? C::NFE : InvShape<C>::NWB;
    const int64_t b = blockIdx.x, x = b % 8, i = b / 8;
    const int64_t w = (i / NW) * 8 + x;
    if (w >= S.total) return;
    const SchurArgs A = seg_schur_args<C>(S, w);
    if constexpr (FB) {
      __shared__ ApplyFbLds<C> lds;
      apply64_fb_body<C>(A, 0, (int)(i % NW), perm, lds);
    } else {
      apply64_body<C>(A, 0, (int)(i % NW));
    }
  }
}



}  // namespace hdgx
